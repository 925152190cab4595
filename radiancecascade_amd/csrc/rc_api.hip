// rc_api.hip -- the C ABI (include/rc_gi.h): context, frame orchestration,
// profiling, parity readbacks, geometry and scene helpers.
//
// A frame is the distance field -- by default the exact transform (an
// occupancy kernel, a column kernel and a row kernel), or 1 seed kernel +
// the global jump-flood passes (+ an optional fused LDS tail) -- then the
// split march, (L - 2) level kernels and 1 level-0/gather kernel, all stream
// ordered on the caller's stream.  The context owns the distance-field
// buffers of its mode, the SDF, the merged levels 1..L-1 and the per-level
// direction tables; it allocates them at create/resize/mode-change time only.
#include <rocprofiler-sdk-roctx/roctx.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <new>
#include <string>
#include <vector>

#include "rc_internal.hpp"
#include "rc_batch.hpp"

using namespace rc;

struct rc_ctx {
    rc_config cfg{};
    int W = 0, H = 0, L = 0;
    Storage storage = Storage::F32;
    DepthConv depth_conv{};
    int sdf_mode = RC_SDF_EXACT;
    uint32_t *seed[2] = {nullptr, nullptr};  // jump flood: seed ping-pong
    uint64_t *edt_bits = nullptr;            // exact transform: occupancy column words
    uint16_t *edt_ends = nullptr;            // exact transform: occupied rows nearest each band
    uint8_t *sdf = nullptr;  // distance codes (rc_common.hpp)
    void *level[RC_MAX_LEVELS] = {};
    size_t level_texels[RC_MAX_LEVELS] = {};
    float *dbg_level0 = nullptr;
    // ray codes of the last frame's in-place marches (RC_FLAG_KEEP_DEBUG),
    // canonical order per level; code_src[i]: where level i's codes are
    // (0 none, 1 the split march's `codes`, 2 `dbg_codes`)
    int *dbg_codes = nullptr;
    size_t dbg_code_off[RC_MAX_LEVELS] = {};
    int code_src[RC_MAX_LEVELS] = {};
    float2 *dirs = nullptr;
    size_t dir_off[RC_MAX_LEVELS] = {};
    int tile[RC_MAX_LEVELS] = {};
    int rpt[RC_MAX_LEVELS] = {};
    // level 1 of the learned frame (marched in place, the network's level 2
    // above): its own tiling, -1 until tuned (then the plain level 1's)
    int surr_tile1 = -1, surr_rpt1 = -1;
    int jfa_fuse = 0;  // largest jump-flood step fused into the LDS tail (0 = none)
    int jfa_rows = 1;  // image rows per wave of a global jump-flood pass
    int mmap[RC_MAX_LEVELS] = {};  // split march: wave-tile order of level i's rays (0 = canonical)
    // split march: levels >= 1 marched by one k_march_all launch into `codes`,
    // then merged top-down; otherwise each level kernel marches its own rays
    int march_split = 1;   // 0 in place, 1 split, 2 split and persistent
    int num_cu = 0;        // compute units of the device (persistent grid size)
    int spec = 1;      // speculation depth of the split march (1, 2, 4)
    int march_m = 1;   // rays per lane, marched back to back (1, 2, 4, 8)
    int *codes = nullptr;
    float4 *emis4 = nullptr;  // the frame's emission interleaved per pixel (plain path)
    size_t code_off[RC_MAX_LEVELS] = {};
    double *norm_ws = nullptr;  // error-norm partials
    float *srgb_table = nullptr;  // composite's 256 sRGB thresholds
    int dbg_seed_buf = -1;
    bool ran = false;
    // host-staging path (rc_update_host / rc_render_host)
    float *g_depth = nullptr, *g_emissive = nullptr, *irr = nullptr;
    bool have_gbuffer = false;
    // whole-frame hipGraph for rc_run_device, keyed on the caller's buffers
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    const void *g_key[3] = {nullptr, nullptr, nullptr};
    // learned far field
    int surr_mode = 0;
    float surr_thr = __builtin_huge_valf();
    MlpDev mlp{};
    bool mlp_ready = false;
    float out_scale = 1.0f;
    uint32_t train_scenes = 0;
    int adam_t = 0;
    float *mlp_partial = nullptr, *mlp_loss_part = nullptr;
    int *mlp_idx = nullptr;
    int mlp_cap_rows = 0;
    bool surr_alloc = false;
    MipArgs mip{};
    ConeSchedule sch{}, sch_far{};
    float4 *cone = nullptr, *cone_far = nullptr;
    uint64_t *occ_bits = nullptr;  // occupancy, one bit per pixel, rows of ceil(W/64) words
    void *surr_a1 = nullptr;   // learned frame's level 1
    void *surr_a2 = nullptr;   // the network's merged level 2
    float *tile_err = nullptr;
    int *flag_count = nullptr, *flag_list = nullptr;
    int *flag_host = nullptr;  // pinned
    int surr_tiles = 0, surr_tiles_x = 0;
    bool surr_ran = false;
    uint8_t *sdf_near = nullptr;   // the learned frame's near distance field (levels 0 and 1)
    // tile fallback: per-level masks of the probes the flagged tiles need, and
    // the device tally of flagged tiles / frames with a fallback
    uint8_t *need = nullptr;
    size_t need_off[RC_MAX_LEVELS] = {};
    size_t need_bytes = 0;
    unsigned long long *tally = nullptr;
    bool last_learned = false;  // the last frame was the learned form (debug readbacks follow it)
    float *scratch_irr = nullptr;  // irradiance of frames run only for their levels
    // the cone schedules (near, far) in device memory
    ConeSchedule *sch_dev = nullptr;
    // profiling
    hipEvent_t ev[2 * RC_NUM_STAGES] = {};
    bool ev_used[RC_NUM_STAGES] = {};
    rc_stats stats{};
    std::string err;
};

static thread_local std::string g_err;

static rc_status fail(rc_ctx *c, rc_status s, const std::string &msg)
{
    g_err = msg;
    if (c) c->err = msg;
    return s;
}

#define HIP_TRY(ctx, expr)                                                                             \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail((ctx), e_ == hipErrorOutOfMemory ? RC_ERR_OOM : RC_ERR_HIP,                    \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                            \
    } while (0)

#define HIP_TRY_ST(expr)                  \
    do {                                  \
        const rc_status st_ = (expr);     \
        if (st_ != RC_OK) return st_;     \
    } while (0)

static bool dims_ok(uint32_t W, uint32_t H) { return W >= 1 && H >= 1 && W <= RC_MAX_DIM && H <= RC_MAX_DIM; }

static uint32_t coverage_levels(uint32_t W, uint32_t H)
{
    const int64_t diag2 = (int64_t)W * W + (int64_t)H * H;
    for (int L = 1; L <= RC_MAX_LEVELS; ++L) {
        const int64_t t = interval_t1(L - 1);
        if (t * t >= diag2) return (uint32_t)L;
    }
    return RC_MAX_LEVELS;
}


static void drop_graph(rc_ctx *c)
{
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->g_key[0] = c->g_key[1] = c->g_key[2] = nullptr;
}

static void release(rc_ctx *c)
{
    drop_graph(c);
    for (auto &p : c->seed) { if (p) hipFree(p); p = nullptr; }
    if (c->edt_bits) hipFree(c->edt_bits);
    if (c->edt_ends) hipFree(c->edt_ends);
    c->edt_bits = nullptr;
    c->edt_ends = nullptr;
    if (c->sdf) hipFree(c->sdf);
    c->sdf = nullptr;
    for (auto &p : c->level) { if (p) hipFree(p); p = nullptr; }
    if (c->dbg_level0) hipFree(c->dbg_level0);
    c->dbg_level0 = nullptr;
    if (c->dbg_codes) hipFree(c->dbg_codes);
    c->dbg_codes = nullptr;
    for (auto &v : c->code_src) v = 0;
    if (c->dirs) hipFree(c->dirs);
    c->dirs = nullptr;
    if (c->codes) hipFree(c->codes);
    c->codes = nullptr;
    if (c->emis4) hipFree(c->emis4);
    c->emis4 = nullptr;
    if (c->norm_ws) hipFree(c->norm_ws);
    c->norm_ws = nullptr;
    if (c->srgb_table) hipFree(c->srgb_table);
    c->srgb_table = nullptr;
    if (c->mip.tex) hipFree(c->mip.tex);
    c->mip.tex = nullptr;
    if (c->occ_bits) hipFree(c->occ_bits);
    c->occ_bits = nullptr;
    if (c->cone) hipFree(c->cone);
    if (c->cone_far) hipFree(c->cone_far);
    c->cone_far = nullptr;
    if (c->sch_dev) hipFree(c->sch_dev);
    c->sch_dev = nullptr;
    if (c->surr_a1) hipFree(c->surr_a1);
    if (c->surr_a2) hipFree(c->surr_a2);
    c->surr_a2 = nullptr;
    if (c->tile_err) hipFree(c->tile_err);
    if (c->flag_count) hipFree(c->flag_count);
    if (c->flag_list) hipFree(c->flag_list);
    if (c->sdf_near) hipFree(c->sdf_near);
    if (c->need) hipFree(c->need);
    if (c->tally) hipFree(c->tally);
    c->sdf_near = nullptr;
    c->need = nullptr;
    c->tally = nullptr;
    c->need_bytes = 0;
    c->cone = nullptr;
    c->surr_a1 = nullptr;
    c->tile_err = nullptr;
    c->flag_count = c->flag_list = nullptr;
    c->surr_alloc = false;
    c->surr_ran = false;
    if (c->scratch_irr) hipFree(c->scratch_irr);
    c->scratch_irr = nullptr;
    if (c->g_depth) hipFree(c->g_depth);
    if (c->g_emissive) hipFree(c->g_emissive);
    if (c->irr) hipFree(c->irr);
    c->g_depth = c->g_emissive = c->irr = nullptr;
    c->have_gbuffer = false;
    c->ran = false;
}

// The distance-field buffers of the context's mode (the other mode's, if
// allocated earlier, are kept until the next resize).
static rc_status ensure_sdf_buffers(rc_ctx *c)
{
    const size_t n = (size_t)c->W * c->H;
    if (c->sdf_mode == RC_SDF_JUMP_FLOOD) {
        for (auto &p : c->seed)
            if (!p) HIP_TRY(c, hipMalloc(&p, n * sizeof(uint32_t)));
    } else {
        if (!c->edt_bits) HIP_TRY(c, hipMalloc(&c->edt_bits, edt_bits_bytes(c->W, c->H)));
        if (!c->edt_ends) HIP_TRY(c, hipMalloc(&c->edt_ends, edt_ends_bytes(c->W, c->H)));
    }
    return RC_OK;
}

static rc_status allocate(rc_ctx *c, uint32_t W, uint32_t H, uint32_t L)
{
    release(c);
    c->W = (int)W;
    c->H = (int)H;
    c->L = (int)L;
    const size_t n = (size_t)W * H;
    HIP_TRY_ST(ensure_sdf_buffers(c));
    HIP_TRY(c, hipMalloc(&c->sdf, sdf_elems((int)W, (int)H)));
    HIP_TRY(c, hipMalloc(&c->emis4, n * sizeof(float4)));
    const size_t tb = texel_bytes(c->storage);
    for (int i = 1; i < (int)L; ++i) {
        const size_t texels = (size_t)grid_dim(W, i) * grid_dim(H, i) * (num_dirs(i) / 4);
        c->level_texels[i] = texels;
        HIP_TRY(c, hipMalloc(&c->level[i], texels * tb));
    }
    c->level_texels[0] = (size_t)grid_dim(W, 0) * grid_dim(H, 0);
    if (c->cfg.flags & RC_FLAG_KEEP_DEBUG) {
        HIP_TRY(c, hipMalloc(&c->dbg_level0, c->level_texels[0] * 16));
        size_t n_codes = 0;
        for (int i = 0; i < (int)L; ++i) {
            c->dbg_code_off[i] = n_codes;
            n_codes += (size_t)grid_dim(W, i) * grid_dim(H, i) * num_dirs(i);
        }
        HIP_TRY(c, hipMalloc(&c->dbg_codes, n_codes * sizeof(int)));
    }
    // direction tables: theta = 2 pi (k + 1/2) / N in double, rounded to fp32
    std::vector<float2> dirs;
    for (int i = 0; i < (int)L; ++i) {
        c->dir_off[i] = dirs.size();
        const int N = num_dirs(i);
        for (int k = 0; k < N; ++k) {
            const double th = 2.0 * M_PI * ((double)k + 0.5) / (double)N;
            dirs.push_back(make_float2((float)cos(th), (float)sin(th)));
        }
        c->tile[i] = default_level_tile(i, (int)L);
        c->rpt[i] = 2;
    }
    c->surr_tile1 = c->surr_rpt1 = -1;
    size_t ncodes = 0;
    for (int i = 1; i < (int)L; ++i) {
        c->code_off[i] = ncodes;
        ncodes += (size_t)grid_dim(W, i) * grid_dim(H, i) * num_dirs(i);
    }
    if (ncodes) HIP_TRY(c, hipMalloc(&c->codes, ncodes * sizeof(int)));
    HIP_TRY(c, hipMalloc(&c->norm_ws, norm_workspace_bytes()));
    HIP_TRY(c, hipMalloc(&c->dirs, dirs.size() * sizeof(float2)));
    HIP_TRY(c, hipMemcpy(c->dirs, dirs.data(), dirs.size() * sizeof(float2), hipMemcpyHostToDevice));
    for (int i = 0; i < RC_MAX_LEVELS; ++i) {
        c->stats.level_rays[i] = i < (int)L ? (uint64_t)grid_dim(W, i) * grid_dim(H, i) * num_dirs(i) : 0;
        c->stats.level_bytes_stored[i] = (i >= 1 && i < (int)L) ? c->level_texels[i] * tb : 0;
    }
    return RC_OK;
}

static LevelArgs level_args(const rc_ctx *c, int i, const float *emissive)
{
    LevelArgs a{};
    a.sdf = c->sdf;
    a.emissive = emissive;
    a.dirs = c->dirs + c->dir_off[i];
    a.upper = (i + 1 < c->L) ? c->level[i + 1] : nullptr;
    a.out = c->level[i];
    a.W = c->W;
    a.H = c->H;
    a.level = i;
    a.PX = grid_dim(c->W, i);
    a.PY = grid_dim(c->H, i);
    a.N = num_dirs(i);
    a.UX = grid_dim(c->W, i + 1);
    a.UY = grid_dim(c->H, i + 1);
    a.t0 = (float)interval_t0(i);
    a.t1 = (float)interval_t1(i);
    a.spacing = (float)probe_spacing(i);
    a.codes = nullptr;
    return a;
}

// The split march merges the top level itself (no level kernel for it) when
// it takes one ray per lane and the top level's lane order keeps each
// texel's 4 directions in one lane quad.
static bool top_merged_in_march(const rc_ctx *c)
{
    return c->L > 1 && c->march_split == 1 && c->march_m == 1 && c->mmap[c->L - 1] != 5;
}

static MarchAllArgs march_all_args(const rc_ctx *c)
{
    MarchAllArgs m{};
    m.sdf = c->sdf;
    m.W = c->W;
    m.H = c->H;
    for (int i = c->L - 1; i >= 1; --i) {
        MarchLevel &l = m.lev[m.nlev++];
        l.dirs = c->dirs + c->dir_off[i];
        l.codes = c->codes + c->code_off[i];
        l.PX = grid_dim(c->W, i);
        l.PY = grid_dim(c->H, i);
        l.N = num_dirs(i);
        l.t0 = (float)interval_t0(i);
        l.t1 = (float)interval_t1(i);
        l.spacing = (float)probe_spacing(i);
        l.log2n = 2 + 2 * i;
        l.inv_px = 1.0f / (float)l.PX;
        march_level_map(l, c->march_split == 1 ? c->mmap[i] : 0);  // the persistent form is canonical
        l.merged = (i == c->L - 1 && top_merged_in_march(c)) ? c->level[i] : nullptr;
        l.storage = (int)c->storage;
        const int64_t rays = (int64_t)l.PX * l.PY * l.N, per = (int64_t)kMarchBlock * c->march_m;
        l.blocks = (int)(((int64_t)l.slots + per - 1) / per);
        l.code_begin = (uint32_t)c->code_off[i];
        l.dir_off = (int)c->dir_off[i];
        m.total_rays += (uint32_t)rays;
    }
    m.codes_base = c->codes;
    m.dirs_base = c->dirs;
    m.emis4 = c->emis4;
    m.n_packets = (m.total_rays + 63) / 64;
    // resident waves: 8 workgroups of 4 waves per CU (the march's occupancy cap)
    const uint32_t cap = (uint32_t)(c->num_cu > 0 ? c->num_cu : 256) * 32u;
    m.n_waves = m.n_packets < cap ? m.n_packets : cap;
    return m;
}

// Stage names of the profiled frame's roctx ranges ("rc:jfa", "rc:level3", ...).
static const char *stage_name(int st)
{
    static const std::vector<std::string> names = [] {
        std::vector<std::string> n(RC_NUM_STAGES, "rc:stage");
        n[RC_STAGE_SEED] = "rc:seed";
        n[RC_STAGE_JFA] = "rc:jfa";
        n[RC_STAGE_JFA_TAIL] = "rc:jfa_tail";
        for (int i = 0; i < RC_MAX_LEVELS; ++i) n[RC_STAGE_LEVEL0 + i] = "rc:level" + std::to_string(i);
        n[RC_STAGE_MARCH] = "rc:march";
        n[RC_STAGE_NEAR_SDF] = "rc:near_sdf";
        n[RC_STAGE_MIP] = "rc:mip";
        n[RC_STAGE_CONE] = "rc:cone";
        n[RC_STAGE_SURROGATE] = "rc:network";
        n[RC_STAGE_FALLBACK] = "rc:fallback";
        return n;
    }();
    return names[st].c_str();
}

// Profiled frames: HIP events around every stage on the frame's stream (the
// stage times of rc_get_stats) and a roctx range with the stage's name (host
// side, for rocprofv3 --marker-trace).
struct StageTimer {
    rc_ctx *c;
    hipStream_t s;
    bool on;
    void begin(int stage)
    {
        if (on) {
            roctxRangePushA(stage_name(stage));
            hipEventRecord(c->ev[2 * stage], s);
            c->ev_used[stage] = true;
        }
    }
    void end(int stage)
    {
        if (on) {
            hipEventRecord(c->ev[2 * stage + 1], s);
            roctxRangePop();
        }
    }
};

static rc_status collect_profile(rc_ctx *c, hipStream_t s, int jfa_launches)
{
    HIP_TRY(c, hipStreamSynchronize(s));
    for (int st = 0; st < RC_NUM_STAGES; ++st) {
        if (!c->ev_used[st]) continue;
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[2 * st], c->ev[2 * st + 1]));
        const uint32_t n = st == RC_STAGE_JFA ? (uint32_t)jfa_launches : 1u;
        if (n == 0) continue;
        c->stats.stage_ms[st] += ms;
        c->stats.stage_launches[st] += n;
    }
    return RC_OK;
}

// The distance field c->sdf (and the interleaved emission c->emis4).  Exact
// mode: the occupancy kernel (stage "seed"), then the band-end and row kernels
// (stage "jfa", 2 launches).  Jump-flood mode: occupancy seeds, then global
// passes down to the fused LDS tail (if any); without a tail the last global
// pass writes the field, and a 1x1 image has no pass.  With `gate` every
// kernel does nothing when *gate == 0 (the learned frame's tile fallback).
static rc_status seeds_and_jfa(rc_ctx *c, const float *depth, const float *emissive, hipStream_t s, bool dbg,
                               const int *gate, StageTimer *tm, int *global_passes)
{
    if (c->sdf_mode == RC_SDF_EXACT) {
        c->dbg_seed_buf = -1;
        if (tm) tm->begin(RC_STAGE_SEED);
        HIP_TRY(c, launch_edt(depth, c->depth_conv, emissive, c->edt_bits, c->edt_ends, c->sdf, c->emis4, c->W, c->H, s,
                              gate, 1));
        if (tm) tm->end(RC_STAGE_SEED);
        if (tm) tm->begin(RC_STAGE_JFA);
        HIP_TRY(c, launch_edt(depth, c->depth_conv, emissive, c->edt_bits, c->edt_ends, c->sdf, c->emis4, c->W, c->H, s,
                              gate, 2));
        if (tm) tm->end(RC_STAGE_JFA);
        if (global_passes) *global_passes = 2;
        return RC_OK;
    }
    if (tm) tm->begin(RC_STAGE_SEED);
    HIP_TRY(c, launch_seed_init(depth, c->depth_conv, emissive, c->seed[0], c->emis4, c->W, c->H, s, gate));
    if (tm) tm->end(RC_STAGE_SEED);
    const int P = jfa_num_passes(c->W, c->H);
    const int tail = jfa_tail_step(c->W, c->H, c->jfa_fuse);
    int cur = 0, passes = 0;
    c->dbg_seed_buf = -1;
    if (tm) tm->begin(RC_STAGE_JFA);
    for (int step = P > 0 ? 1 << (P - 1) : 0; step >= 1 && step > tail; step >>= 1) {
        const bool final_pass = tail == 0 && step == 1;
        uint32_t *dst = final_pass ? (dbg ? c->seed[cur ^ 1] : nullptr) : c->seed[cur ^ 1];
        HIP_TRY(c, launch_jfa_pass(c->seed[cur], dst, c->sdf, c->W, c->H, step, final_pass, s, gate, c->jfa_rows));
        if (final_pass) c->dbg_seed_buf = dbg ? (cur ^ 1) : -1;
        else cur ^= 1;
        ++passes;
    }
    if (tm) tm->end(RC_STAGE_JFA);
    if (tail > 0 || P == 0) {
        if (tm) tm->begin(RC_STAGE_JFA_TAIL);
        HIP_TRY(c, launch_jfa_tail(c->seed[cur], dbg ? c->seed[cur ^ 1] : nullptr, c->sdf, c->W, c->H, tail, s, gate));
        if (tm) tm->end(RC_STAGE_JFA_TAIL);
        c->dbg_seed_buf = dbg ? (cur ^ 1) : -1;
    }
    if (global_passes) *global_passes = passes;
    return RC_OK;
}

static rc_status run_frame(rc_ctx *c, const float *depth, const float *emissive, float *irr, hipStream_t s,
                           bool quiet = false)
{
    const bool prof = !quiet && (c->cfg.flags & RC_FLAG_PROFILE) != 0;
    const bool dbg = (c->cfg.flags & RC_FLAG_KEEP_DEBUG) != 0;
    StageTimer tm{c, s, prof};
    for (auto &u : c->ev_used) u = false;

    int global_passes = 0;
    rc_status js = seeds_and_jfa(c, depth, emissive, s, dbg, nullptr, &tm, &global_passes);
    if (js != RC_OK) return js;

    const bool split = c->march_split && c->L > 1;
    if (split) {
        tm.begin(RC_STAGE_MARCH);
        if (c->march_split == 2) HIP_TRY(c, launch_march_persist(march_all_args(c), c->spec, s));
        else HIP_TRY(c, launch_march_all(march_all_args(c), c->spec, c->march_m, s));
        tm.end(RC_STAGE_MARCH);
    }
    const int top = split && top_merged_in_march(c) ? c->L - 2 : c->L - 1;  // the march merged level L-1
    for (int i = top; i >= 1; --i) {
        tm.begin(RC_STAGE_LEVEL0 + i);
        LevelArgs a = level_args(c, i, emissive);
        a.emis4 = c->emis4;
        if (split) a.codes = c->codes + c->code_off[i];
        else if (dbg) a.dbg_codes = c->dbg_codes + c->dbg_code_off[i];
        HIP_TRY(c, launch_level(a, c->storage, c->tile[i], c->rpt[i], s));
        tm.end(RC_STAGE_LEVEL0 + i);
        c->code_src[i] = split ? 1 : (dbg ? 2 : 0);
    }
    if (top < c->L - 1) c->code_src[c->L - 1] = 1;  // the march stored the top level's codes too
    tm.begin(RC_STAGE_LEVEL0);
    LevelArgs a0 = level_args(c, 0, emissive);
    if (dbg) a0.dbg_codes = c->dbg_codes + c->dbg_code_off[0];
    HIP_TRY(c, launch_level0_gather(a0, c->storage, irr, dbg ? c->dbg_level0 : nullptr, s));
    tm.end(RC_STAGE_LEVEL0);
    c->code_src[0] = dbg ? 2 : 0;

    if (prof) {
        rc_status st = collect_profile(c, s, global_passes);
        if (st != RC_OK) return st;
    }
    if (!quiet) {
        c->stats.runs += 1;
        c->ran = true;
        c->last_learned = false;
    }
    return RC_OK;
}

// ---- learned far field ---------------------------------------------------------
static void free_mlp(rc_ctx *c)
{
    void *p[] = {c->mlp.master, c->mlp.adam_m, c->mlp.adam_v, c->mlp.wt, c->mlp.wk, c->mlp.bias, c->mlp.in_mean,
                 c->mlp.in_istd, c->mlp_partial, c->mlp_loss_part, c->mlp_idx};
    for (void *q : p)
        if (q) hipFree(q);
    c->mlp = MlpDev{};
    c->mlp_partial = c->mlp_loss_part = nullptr;
    c->mlp_idx = nullptr;
    c->mlp_cap_rows = 0;
    if (c->flag_host) hipHostFree(c->flag_host);
    c->flag_host = nullptr;
}

static rc_status alloc_mlp(rc_ctx *c)
{
    HIP_TRY(c, hipMalloc(&c->mlp.master, kMlpParams * sizeof(float)));
    HIP_TRY(c, hipMalloc(&c->mlp.adam_m, kMlpParams * sizeof(float)));
    HIP_TRY(c, hipMalloc(&c->mlp.adam_v, kMlpParams * sizeof(float)));
    HIP_TRY(c, hipMalloc(&c->mlp.wt, kMlpWt * sizeof(uint16_t)));
    HIP_TRY(c, hipMalloc(&c->mlp.wk, kMlpWt * sizeof(uint16_t)));
    HIP_TRY(c, hipMalloc(&c->mlp.bias, 4 * 64 * sizeof(float)));
    HIP_TRY(c, hipMalloc(&c->mlp.in_mean, kMlpIn * sizeof(float)));
    HIP_TRY(c, hipMalloc(&c->mlp.in_istd, kMlpIn * sizeof(float)));
    HIP_TRY(c, hipMemset(c->mlp.master, 0, kMlpParams * sizeof(float)));
    HIP_TRY(c, hipMemset(c->mlp.adam_m, 0, kMlpParams * sizeof(float)));
    HIP_TRY(c, hipMemset(c->mlp.adam_v, 0, kMlpParams * sizeof(float)));
    HIP_TRY(c, hipHostMalloc(&c->flag_host, sizeof(int), hipHostMallocDefault));
    return RC_OK;
}

// Frame buffers of the learned far field (allocated on first use, freed with
// the frame buffers on resize).
static rc_status ensure_surrogate(rc_ctx *c)
{
    if (c->surr_alloc) return RC_OK;
    if (c->L < 3) return fail(c, RC_ERR_INVALID, "learned far field needs levels >= 3");
    const int top = cone_top_level(c->W, c->H);
    mip_layout(c->W, c->H, top, c->mip);
    cone_schedule(c->W, c->H, c->mip, c->sch, c->sch_far);
    const size_t cells = mip_cells(c->mip);
    if (cells * sizeof(float4) > 0x7FFFFFFFull) return fail(c, RC_ERR_INVALID, "learned far field: image too large");
    HIP_TRY(c, hipMalloc(&c->mip.tex, cells * sizeof(uint4)));
    HIP_TRY(c, hipMemset(c->mip.tex, 0, cells * sizeof(uint4)));  // the zero borders
    HIP_TRY(c, hipMalloc(&c->occ_bits, (size_t)c->H * ((c->W + 63) / 64) * sizeof(uint64_t)));
    const int QX = grid_dim(c->W, 2), QY = grid_dim(c->H, 2);
    HIP_TRY(c, hipMalloc(&c->cone, (size_t)QX * QY * 16 * sizeof(float4)));
    HIP_TRY(c, hipMalloc(&c->cone_far, (size_t)grid_dim(c->W, 3) * grid_dim(c->H, 3) * 16 * sizeof(float4)));
    const int PX = grid_dim(c->W, 1), PY = grid_dim(c->H, 1);
    HIP_TRY(c, hipMalloc(&c->surr_a1, (size_t)PX * PY * 4 * texel_bytes(c->storage)));
    HIP_TRY(c, hipMalloc(&c->surr_a2, (size_t)QX * QY * 16 * texel_bytes(c->storage)));
    c->surr_tiles_x = (QX + 7) / 8;  // tiles of 8x8 level-2 probes
    c->surr_tiles = c->surr_tiles_x * ((QY + 7) / 8);
    HIP_TRY(c, hipMalloc(&c->tile_err, (size_t)c->surr_tiles * sizeof(float)));
    HIP_TRY(c, hipMalloc(&c->flag_list, (size_t)c->surr_tiles * sizeof(int)));
    HIP_TRY(c, hipMalloc(&c->flag_count, sizeof(int)));
    HIP_TRY(c, hipMemset(c->flag_count, 0, sizeof(int)));
    HIP_TRY(c, hipMalloc(&c->sdf_near, sdf_elems(c->W, c->H)));
    c->need_bytes = 0;
    for (int i = 1; i < c->L; ++i) {
        c->need_off[i] = c->need_bytes;
        c->need_bytes += (size_t)grid_dim(c->W, i) * grid_dim(c->H, i);
    }
    HIP_TRY(c, hipMalloc(&c->need, c->need_bytes));
    HIP_TRY(c, hipMalloc(&c->tally, 2 * sizeof(unsigned long long)));
    HIP_TRY(c, hipMemset(c->tally, 0, 2 * sizeof(unsigned long long)));
    HIP_TRY(c, hipMalloc(&c->sch_dev, 2 * sizeof(ConeSchedule)));
    HIP_TRY(c, hipMemcpy(c->sch_dev, &c->sch, sizeof(ConeSchedule), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->sch_dev + 1, &c->sch_far, sizeof(ConeSchedule), hipMemcpyHostToDevice));
    c->surr_alloc = true;
    return RC_OK;
}

// Cone launches: the far segment on the level-3 grid; the near segment on
// the level-2 grid, appending the far one when `append` (else the network
// kernel appends it).
static ConeArgs cone_args_far(rc_ctx *c)
{
    ConeArgs a{};
    a.mip = c->mip;
    a.mip_bytes = (uint32_t)(mip_cells(c->mip) * sizeof(uint4));
    a.dirs = c->dirs + c->dir_off[1];
    a.sch = c->sch_dev + 1;
    a.cone = c->cone_far;
    a.QX = grid_dim(c->W, 3);
    a.QY = grid_dim(c->H, 3);
    a.spacing = (float)probe_spacing(3);
    a.far = nullptr;
    return a;
}

static ConeArgs cone_args_near(rc_ctx *c, bool append)
{
    ConeArgs a = cone_args_far(c);
    a.sch = c->sch_dev;
    a.cone = c->cone;
    a.QX = grid_dim(c->W, 2);
    a.QY = grid_dim(c->H, 2);
    a.spacing = (float)probe_spacing(2);
    a.far = append ? c->cone_far : nullptr;
    a.FX = grid_dim(c->W, 3);
    a.FY = grid_dim(c->H, 3);
    return a;
}

static rc_status launch_cones(rc_ctx *c, hipStream_t s)
{
    HIP_TRY(c, launch_cone(cone_args_far(c), s));
    HIP_TRY(c, launch_cone(cone_args_near(c, true), s));
    return RC_OK;
}

// occupancy bits + mip 1..2 -> mip 3..6 + near distance field -> mip
// 7..top -> near + far cones -> network (appending the far segment) ->
// level 1 (plain kernel on the near field) -> level 0 + gather.
static rc_status run_frame_surrogate(rc_ctx *c, const float *depth, const float *emissive, float *irr, hipStream_t s,
                                     bool quiet)
{
    const bool prof = !quiet && (c->cfg.flags & RC_FLAG_PROFILE) != 0;
    const bool dbg = (c->cfg.flags & RC_FLAG_KEEP_DEBUG) != 0;
    StageTimer tm{c, s, prof};
    for (auto &u : c->ev_used) u = false;
    tm.begin(RC_STAGE_SEED);
    HIP_TRY(c, launch_seed_surr(depth, c->depth_conv, emissive, c->occ_bits, c->mip, c->W, c->H, s));
    tm.end(RC_STAGE_SEED);
    // One stream: work that could overlap shares a launch instead (kernels
    // on several captured streams wait microseconds for each other across
    // hardware queues).  A profiled frame runs the stages one by one.
    if (!prof) {
        // the top mip levels (one workgroup) beside the near distance field;
        // near and far cones in one launch (the network kernel appends the
        // far segment)
        HIP_TRY(c, launch_mip_top_near(c->mip, c->flag_count, c->occ_bits, c->sdf_near, c->W, c->H, s));
        HIP_TRY(c, launch_cone_pair(cone_args_near(c, false), cone_args_far(c), s));
    } else {
        tm.begin(RC_STAGE_NEAR_SDF);
        HIP_TRY(c, launch_near_sdf(c->occ_bits, c->sdf_near, c->W, c->H, s));
        tm.end(RC_STAGE_NEAR_SDF);
        tm.begin(RC_STAGE_MIP);
        HIP_TRY(c, launch_mip_top(c->mip, c->flag_count, s));
        tm.end(RC_STAGE_MIP);
        tm.begin(RC_STAGE_CONE);
        HIP_TRY(c, launch_cone(cone_args_far(c), s));
        HIP_TRY(c, launch_cone(cone_args_near(c, false), s));
        tm.end(RC_STAGE_CONE);
    }
    SurrMlpArgs a{};
    a.cone = c->cone;
    a.far = c->cone_far;
    a.FX = grid_dim(c->W, 3);
    a.FY = grid_dim(c->H, 3);
    a.cone_out = dbg ? c->cone : nullptr;
    a.out = c->surr_a2;
    a.tile_err = c->tile_err;
    a.flag_count = c->flag_count;
    a.flag_list = c->flag_list;
    a.QX = grid_dim(c->W, 2);
    a.QY = grid_dim(c->H, 2);
    a.tiles_x = c->surr_tiles_x;
    a.tiles = c->surr_tiles;
    a.out_scale = c->out_scale;
    a.threshold = c->surr_thr;
    tm.begin(RC_STAGE_SURROGATE);
    HIP_TRY(c, launch_surr_mlp(a, c->mlp, c->storage, s));
    tm.end(RC_STAGE_SURROGATE);
    // Tile fallback (finite threshold): the plain far field for the flagged
    // tiles only, decided on the device.  Every kernel reads the flagged-tile
    // count the network kernel left and does nothing when it is zero; the
    // masked level kernels compute only the probes the flagged tiles need
    // (their bilinear neighbourhood at each level up), marched in place on
    // the whole frame's jump-flood field.
    const bool fallback = c->surr_thr < __builtin_huge_valf();
    if (fallback) {
        tm.begin(RC_STAGE_FALLBACK);
        HIP_TRY_ST(seeds_and_jfa(c, depth, emissive, s, false, c->flag_count, nullptr, nullptr));
        FallbackMaskArgs fm{};
        fm.count = c->flag_count;
        fm.list = c->flag_list;
        fm.need = c->need;
        for (int i = 1; i < c->L; ++i) {
            fm.need_off[i] = c->need_off[i];
            fm.PX[i] = grid_dim(c->W, i);
            fm.PY[i] = grid_dim(c->H, i);
        }
        fm.L = c->L;
        fm.tiles_x = c->surr_tiles_x;
        fm.max_tiles = c->surr_tiles;
        fm.need_bytes = c->need_bytes;
        HIP_TRY(c, launch_fallback_masks(fm, s));
        for (int i = c->L - 1; i >= 2; --i) {
            LevelArgs ai = level_args(c, i, emissive);
            ai.emis4 = c->emis4;
            ai.gate = c->flag_count;
            ai.need = c->need + c->need_off[i];
            HIP_TRY(c, launch_level(ai, c->storage, c->tile[i], c->rpt[i], s));
        }
        tm.end(RC_STAGE_FALLBACK);
    }
    // level 1: the plain level kernel, marching in place on the near field,
    // merging the network's level 2
    tm.begin(RC_STAGE_LEVEL0 + 1);
    LevelArgs a1 = level_args(c, 1, emissive);
    a1.sdf = c->sdf_near;
    a1.upper = c->surr_a2;
    a1.out = c->surr_a1;
    a1.codes = nullptr;
    if (dbg) a1.dbg_codes = c->dbg_codes + c->dbg_code_off[1];
    HIP_TRY(c, launch_level(a1, c->storage, c->surr_tile1 >= 0 ? c->surr_tile1 : c->tile[1],
                            c->surr_rpt1 >= 0 ? c->surr_rpt1 : c->rpt[1], s));
    if (fallback) {
        // the flagged tiles' level 1 again, plain: marched on the jump-flood
        // field and merged from the plain level 2, over the learned one
        LevelArgs af = level_args(c, 1, emissive);
        af.emis4 = c->emis4;
        af.out = c->surr_a1;
        af.gate = c->flag_count;
        af.need = c->need + c->need_off[1];
        HIP_TRY(c, launch_level(af, c->storage, c->tile[1], c->rpt[1], s));
    }
    tm.end(RC_STAGE_LEVEL0 + 1);
    tm.begin(RC_STAGE_LEVEL0);
    LevelArgs a0 = level_args(c, 0, emissive);
    a0.sdf = c->sdf_near;
    a0.upper = c->surr_a1;
    if (dbg) a0.dbg_codes = c->dbg_codes + c->dbg_code_off[0];
    HIP_TRY(c, launch_level0_gather(a0, c->storage, irr, dbg ? c->dbg_level0 : nullptr, s));
    tm.end(RC_STAGE_LEVEL0);
    if (fallback) HIP_TRY(c, launch_fallback_tally(c->flag_count, c->tally, s));
    for (int i = 0; i < RC_MAX_LEVELS; ++i) c->code_src[i] = (dbg && i <= 1) ? 2 : 0;
    if (prof) {
        rc_status st = collect_profile(c, s, 0);
        if (st != RC_OK) return st;
    }
    if (!quiet) {
        c->stats.runs += 1;
        c->ran = true;
    }
    c->surr_ran = true;
    c->last_learned = true;
    return RC_OK;
}

static rc_status run_frame_any(rc_ctx *c, const float *depth, const float *emissive, float *irr, hipStream_t s,
                               bool quiet)
{
    if (c->surr_mode) return run_frame_surrogate(c, depth, emissive, irr, s, quiet);
    return run_frame(c, depth, emissive, irr, s, quiet);
}

// With RC_FLAG_GRAPH the frame's launches are replayed as one graph;
// profiling and debug runs, which record events or extra outputs, always
// launch directly.
static bool graph_enabled(const rc_ctx *c)
{
    return (c->cfg.flags & RC_FLAG_GRAPH) && !(c->cfg.flags & (RC_FLAG_PROFILE | RC_FLAG_KEEP_DEBUG));
}

static rc_status run_frame_graph(rc_ctx *c, const float *depth, const float *emissive, float *irr, hipStream_t s)
{
    if (!c->gexec || c->g_key[0] != depth || c->g_key[1] != emissive || c->g_key[2] != irr) {
        drop_graph(c);
        if (!c->cap_stream) HIP_TRY(c, hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
        HIP_TRY(c, hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
        rc_status st = run_frame_any(c, depth, emissive, irr, c->cap_stream, true);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->cap_stream, &g);
        if (st != RC_OK) {
            if (g) hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return fail(c, RC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->gexec = nullptr;
            return fail(c, RC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        c->g_key[0] = depth;
        c->g_key[1] = emissive;
        c->g_key[2] = irr;
    }
    HIP_TRY(c, hipGraphLaunch(c->gexec, s));
    c->stats.runs += 1;
    c->ran = true;
    return RC_OK;
}

// One frame on stream s.  In learned mode with a finite threshold the tile
// fallback is part of the frame's launches and decided on the device (no
// host synchronisation); rc_get_stats reads its tally.
static rc_status frame(rc_ctx *c, const float *depth, const float *emissive, float *irr, hipStream_t s)
{
    if (c->surr_mode) {
        if (!c->mlp_ready)
            return fail(c, RC_ERR_STATE, "learned far field: no network (rc_mlp_set, rc_surrogate_load or rc_surrogate_train)");
        rc_status st = ensure_surrogate(c);
        if (st != RC_OK) return st;
    }
    rc_status st = graph_enabled(c) ? run_frame_graph(c, depth, emissive, irr, s)
                                    : run_frame_any(c, depth, emissive, irr, s, false);
    if (st != RC_OK || !c->surr_mode) return st;
    c->stats.surrogate_frames += 1;
    c->stats.tiles_per_frame = (uint64_t)c->surr_tiles;
    return RC_OK;
}

// Training rows of the G-buffer at the context's size: the plain frame, then
// the mip pyramid and cones of the same G-buffer.
static rc_status gen_samples(rc_ctx *c, const float *depth, const float *emissive, float *feats, float *targets,
                             hipStream_t s)
{
    rc_status st = ensure_surrogate(c);
    if (st != RC_OK) return st;
    if (!c->scratch_irr) HIP_TRY(c, hipMalloc(&c->scratch_irr, 3 * (size_t)c->W * c->H * sizeof(float)));
    st = run_frame(c, depth, emissive, c->scratch_irr, s, true);
    if (st != RC_OK) return st;
    HIP_TRY(c, launch_seed_surr(depth, c->depth_conv, emissive, c->occ_bits, c->mip, c->W, c->H, s));
    HIP_TRY(c, launch_mip_top(c->mip, c->flag_count, s));
    st = launch_cones(c, s);
    if (st != RC_OK) return st;
    SurrSampleArgs a{};
    a.cone = c->cone;
    a.upper = c->level[2];
    a.feats = feats;
    a.targets = targets;
    a.QX = grid_dim(c->W, 2);
    a.QY = grid_dim(c->H, 2);
    HIP_TRY(c, launch_surr_sample(a, c->storage, s));
    return RC_OK;
}

static uint64_t mix64(uint64_t &state)
{
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Scratch for the gradient hook: per-workgroup partial gradients and losses.
static rc_status ensure_mlp_scratch(rc_ctx *c, int parts, int rows)
{
    if (parts > c->mlp_cap_rows) {
        if (c->mlp_partial) hipFree(c->mlp_partial);
        if (c->mlp_loss_part) hipFree(c->mlp_loss_part);
        c->mlp_partial = c->mlp_loss_part = nullptr;
        HIP_TRY(c, hipMalloc(&c->mlp_partial, (size_t)parts * kMlpParams * sizeof(float)));
        HIP_TRY(c, hipMalloc(&c->mlp_loss_part, (size_t)parts * sizeof(float)));
        c->mlp_cap_rows = parts;
    }
    (void)rows;
    return RC_OK;
}

static rc_status mlp_upload(rc_ctx *c, const float *params, const float *mean, const float *istd, float out_scale)
{
    HIP_TRY(c, hipMemcpy(c->mlp.master, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->mlp.in_mean, mean, kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->mlp.in_istd, istd, kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c, launch_mlp_refresh(c->mlp, nullptr));
    HIP_TRY(c, hipMemset(c->mlp.adam_m, 0, kMlpParams * sizeof(float)));
    HIP_TRY(c, hipMemset(c->mlp.adam_v, 0, kMlpParams * sizeof(float)));
    HIP_TRY(c, hipDeviceSynchronize());
    c->adam_t = 0;
    c->out_scale = out_scale;
    c->mlp_ready = true;
    drop_graph(c);
    return RC_OK;
}

// ---------------------------------------------------------------------------
extern "C" {

uint32_t rc_abi_version(void) { return RC_ABI_VERSION; }

const char *rc_status_string(rc_status s)
{
    switch (s) {
    case RC_OK: return "ok";
    case RC_ERR_INVALID: return "invalid argument";
    case RC_ERR_NULL: return "null pointer";
    case RC_ERR_NO_DEVICE: return "no HIP device";
    case RC_ERR_HIP: return "HIP runtime error";
    case RC_ERR_OOM: return "device out of memory";
    case RC_ERR_STATE: return "call out of order";
    }
    return "unknown status";
}

const char *rc_last_error(const rc_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

rc_status rc_device_count(int *count)
{
    if (!count) return fail(nullptr, RC_ERR_NULL, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        *count = 0;
        return fail(nullptr, RC_ERR_NO_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = n;
    return RC_OK;
}

rc_status rc_create(const rc_config *cfg, rc_ctx **out)
{
    if (!cfg || !out) return fail(nullptr, RC_ERR_NULL, "rc_create: cfg or out is NULL");
    *out = nullptr;
    if (cfg->version != RC_ABI_VERSION) return fail(nullptr, RC_ERR_INVALID, "rc_create: ABI version mismatch");
    if (!dims_ok(cfg->width, cfg->height)) return fail(nullptr, RC_ERR_INVALID, "rc_create: width/height out of range");
    if (cfg->levels > RC_MAX_LEVELS) return fail(nullptr, RC_ERR_INVALID, "rc_create: levels out of range");
    if (cfg->storage > RC_STORAGE_BF16) return fail(nullptr, RC_ERR_INVALID, "rc_create: unknown storage");
    if (cfg->flags & ~(RC_FLAG_PROFILE | RC_FLAG_KEEP_DEBUG | RC_FLAG_GRAPH)) return fail(nullptr, RC_ERR_INVALID, "rc_create: unknown flags");
    if (cfg->depth_mode > RC_DEPTH_NDC) return fail(nullptr, RC_ERR_INVALID, "rc_create: unknown depth_mode");
    if (cfg->depth_mode == RC_DEPTH_NDC &&
        !(cfg->z_near > 0.0f && cfg->z_far > cfg->z_near && cfg->z_far < __builtin_huge_valf()))
        return fail(nullptr, RC_ERR_INVALID, "rc_create: NDC depth needs 0 < z_near < z_far < inf");
    int ndev = 0;
    if (rc_device_count(&ndev) != RC_OK) return RC_ERR_NO_DEVICE;
    if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, RC_ERR_INVALID, "rc_create: device ordinal out of range");
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return fail(nullptr, RC_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));

    rc_ctx *c = new (std::nothrow) rc_ctx();
    if (!c) return fail(nullptr, RC_ERR_OOM, "rc_create: host allocation failed");
    if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, cfg->device) != hipSuccess)
        c->num_cu = 0;
    c->cfg = *cfg;
    c->storage = (Storage)cfg->storage;
    if (cfg->depth_mode == RC_DEPTH_NDC) {
        // fp32 constants of D = (n f) / (f - d (f - n)), each rounded once
        const float n = cfg->z_near, f = cfg->z_far;
        c->depth_conv = DepthConv{1, n * f, f, f - n};
    }
    const uint32_t L = cfg->levels ? cfg->levels : coverage_levels(cfg->width, cfg->height);
    c->cfg.levels = L;
    for (auto &ev : c->ev) {
        if (hipEventCreate(&ev) != hipSuccess) { rc_destroy(c); return fail(nullptr, RC_ERR_HIP, "hipEventCreate failed"); }
    }
    rc_status st = alloc_mlp(c);
    if (st == RC_OK) st = allocate(c, cfg->width, cfg->height, L);
    if (st != RC_OK) {
        std::string msg = c->err;
        rc_destroy(c);
        return fail(nullptr, st, msg);
    }
    c->stats.version = RC_ABI_VERSION;
    *out = c;
    return RC_OK;
}

rc_status rc_resize(rc_ctx *c, uint32_t W, uint32_t H, uint32_t L)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_resize: ctx is NULL");
    if (!dims_ok(W, H) || L > RC_MAX_LEVELS) return fail(c, RC_ERR_INVALID, "rc_resize: dimensions out of range");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipDeviceSynchronize());
    if (L == 0) L = coverage_levels(W, H);
    c->cfg.width = W;
    c->cfg.height = H;
    c->cfg.levels = L;
    rc_status st = allocate(c, W, H, L);
    if (st == RC_OK && c->surr_mode) st = ensure_surrogate(c);
    return st;
}

rc_status rc_run_device(rc_ctx *c, const rc_gbuffer *g, float *irr, void *stream)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_run_device: ctx is NULL");
    if (!g || !g->depth || !g->emissive || !irr) return fail(c, RC_ERR_NULL, "rc_run_device: G-buffer or output is NULL");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    return frame(c, g->depth, g->emissive, irr, (hipStream_t)stream);
}

rc_status rc_update_host(rc_ctx *c, const rc_gbuffer *host)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_update_host: ctx is NULL");
    if (!host || !host->depth || !host->emissive) return fail(c, RC_ERR_NULL, "rc_update_host: G-buffer is NULL");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    const size_t n = (size_t)c->W * c->H;
    if (!c->g_depth) {
        HIP_TRY(c, hipMalloc(&c->g_depth, n * sizeof(float)));
        HIP_TRY(c, hipMalloc(&c->g_emissive, 3 * n * sizeof(float)));
        HIP_TRY(c, hipMalloc(&c->irr, 3 * n * sizeof(float)));
    }
    HIP_TRY(c, hipMemcpy(c->g_depth, host->depth, n * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->g_emissive, host->emissive, 3 * n * sizeof(float), hipMemcpyHostToDevice));
    c->have_gbuffer = true;
    return RC_OK;
}

rc_status rc_render_host(rc_ctx *c, float *irr_host)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_render_host: ctx is NULL");
    if (!irr_host) return fail(c, RC_ERR_NULL, "rc_render_host: output is NULL");
    if (!c->have_gbuffer) return fail(c, RC_ERR_STATE, "rc_render_host: no G-buffer uploaded (call rc_update_host)");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    rc_status st = frame(c, c->g_depth, c->g_emissive, c->irr, nullptr);
    if (st != RC_OK) return st;
    HIP_TRY(c, hipMemcpy(irr_host, c->irr, 3 * (size_t)c->W * c->H * sizeof(float), hipMemcpyDeviceToHost));
    return RC_OK;
}

rc_status rc_host_alloc(uint64_t bytes, void **out)
{
    if (!out) return fail(nullptr, RC_ERR_NULL, "rc_host_alloc: out is NULL");
    *out = nullptr;
    if (bytes == 0) return fail(nullptr, RC_ERR_INVALID, "rc_host_alloc: zero bytes");
    const hipError_t e = hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(nullptr, e == hipErrorOutOfMemory ? RC_ERR_OOM : RC_ERR_HIP,
                    std::string("rc_host_alloc: ") + hipGetErrorString(e));
    }
    return RC_OK;
}

void rc_host_free(void *ptr)
{
    if (ptr) hipHostFree(ptr);
}

void rc_destroy(rc_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->cfg.device);
    hipDeviceSynchronize();
    release(c);
    free_mlp(c);
    for (auto &ev : c->ev)
        if (ev) hipEventDestroy(ev);
    if (c->cap_stream) hipStreamDestroy(c->cap_stream);
    delete c;
}

rc_status rc_get_config(const rc_ctx *c, rc_config *out)
{
    if (!c || !out) return fail(nullptr, RC_ERR_NULL, "rc_get_config: NULL argument");
    *out = c->cfg;
    return RC_OK;
}

rc_status rc_get_stats(const rc_ctx *c, rc_stats *out)
{
    if (!c || !out) return fail(nullptr, RC_ERR_NULL, "rc_get_stats: NULL argument");
    *out = c->stats;
    if (c->tally) {
        // the tile fallback's device tally (waits for the frames enqueued so far)
        unsigned long long t[2] = {0, 0};
        hipError_t e = hipSetDevice(c->cfg.device);
        if (e == hipSuccess) e = hipMemcpy(t, c->tally, sizeof(t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(nullptr, RC_ERR_HIP, std::string("rc_get_stats: ") + hipGetErrorString(e));
        out->flagged_tiles += t[0];
        out->fallback_frames += t[1];
    }
    return RC_OK;
}

rc_status rc_reset_stats(rc_ctx *c)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_reset_stats: ctx is NULL");
    for (auto &v : c->stats.stage_ms) v = 0.0;
    for (auto &v : c->stats.stage_launches) v = 0;
    c->stats.runs = 0;
    c->stats.surrogate_frames = c->stats.fallback_frames = c->stats.flagged_tiles = 0;
    if (c->tally) {
        HIP_TRY(c, hipSetDevice(c->cfg.device));
        HIP_TRY(c, hipDeviceSynchronize());
        HIP_TRY(c, hipMemset(c->tally, 0, 2 * sizeof(unsigned long long)));
    }
    return RC_OK;
}

rc_status rc_debug_seeds(rc_ctx *c, uint32_t *dev_out, void *stream)
{
    if (!c || !dev_out) return fail(c, RC_ERR_NULL, "rc_debug_seeds: NULL argument");
    if (c->sdf_mode != RC_SDF_JUMP_FLOOD)
        return fail(c, RC_ERR_STATE, "rc_debug_seeds: the seed field exists in the jump-flood mode only");
    if (!c->ran || c->dbg_seed_buf < 0) return fail(c, RC_ERR_STATE, "rc_debug_seeds: needs a run with RC_FLAG_KEEP_DEBUG");
    HIP_TRY(c, hipMemcpyAsync(dev_out, c->seed[c->dbg_seed_buf], (size_t)c->W * c->H * 4, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
    return RC_OK;
}

rc_status rc_debug_sdf(rc_ctx *c, float *dev_out, void *stream)
{
    if (!c || !dev_out) return fail(c, RC_ERR_NULL, "rc_debug_sdf: NULL argument");
    if (!c->ran) return fail(c, RC_ERR_STATE, "rc_debug_sdf: no run yet");
    // after a learned frame: the near field its levels 0 and 1 marched on
    HIP_TRY(c, launch_sdf_untile(c->last_learned ? c->sdf_near : c->sdf, dev_out, c->W, c->H, (hipStream_t)stream));
    return RC_OK;
}

rc_status rc_debug_level(rc_ctx *c, uint32_t level, float *dev_out, void *stream)
{
    if (!c || !dev_out) return fail(c, RC_ERR_NULL, "rc_debug_level: NULL argument");
    if ((int)level >= c->L) return fail(c, RC_ERR_INVALID, "rc_debug_level: level out of range");
    if (!c->ran) return fail(c, RC_ERR_STATE, "rc_debug_level: no run yet");
    if (level == 0) {
        if (!c->dbg_level0) return fail(c, RC_ERR_STATE, "rc_debug_level: level 0 needs RC_FLAG_KEEP_DEBUG");
        HIP_TRY(c, hipMemcpyAsync(dev_out, c->dbg_level0, c->level_texels[0] * 16, hipMemcpyDeviceToDevice,
                                  (hipStream_t)stream));
        return RC_OK;
    }
    const void *src = c->level[level];
    if (c->last_learned) {
        // the learned frame's own levels: the network's level 2 and level 1
        // merged from it; it computes no plain level above 2
        if (level > 2) return fail(c, RC_ERR_STATE, "rc_debug_level: the learned frame has no level above 2");
        src = level == 2 ? c->surr_a2 : c->surr_a1;
    }
    HIP_TRY(c, launch_level_to_f32(src, c->storage, c->level_texels[level], dev_out, (hipStream_t)stream));
    return RC_OK;
}

rc_status rc_debug_march(rc_ctx *c, uint32_t level, int32_t *hit_dev, int64_t *index_dev, void *stream)
{
    if (!c || !hit_dev) return fail(c, RC_ERR_NULL, "rc_debug_march: NULL argument");
    if ((int)level >= c->L) return fail(c, RC_ERR_INVALID, "rc_debug_march: level out of range");
    if (!c->ran) return fail(c, RC_ERR_STATE, "rc_debug_march: no run yet");
    const hipStream_t s = (hipStream_t)stream;
    const size_t rays = (size_t)grid_dim(c->W, (int)level) * grid_dim(c->H, (int)level) * num_dirs((int)level);
    // the codes the last frame's kernels stored: the split march's buffer,
    // or the in-place marches' debug copy
    const int src = c->code_src[level];
    if (src == 1) {
        HIP_TRY(c, hipMemcpyAsync(hit_dev, c->codes + c->code_off[level], rays * sizeof(int), hipMemcpyDeviceToDevice, s));
    } else if (src == 2) {
        HIP_TRY(c, hipMemcpyAsync(hit_dev, c->dbg_codes + c->dbg_code_off[level], rays * sizeof(int),
                                  hipMemcpyDeviceToDevice, s));
    } else {
        return fail(c, RC_ERR_STATE, "rc_debug_march: this level's rays were not marched in the last frame with "
                                     "RC_FLAG_KEEP_DEBUG (or, in the split march, at all)");
    }
    if (index_dev) {
        if (level == 0) return fail(c, RC_ERR_INVALID, "rc_debug_march: level 0 has no split-march decode");
        const MarchAllArgs m = march_all_args(c);
        HIP_TRY(c, hipMemsetAsync(index_dev, 0xFF, rays * sizeof(int64_t), s));  // -1: no slot
        HIP_TRY(c, launch_ray_decode(m.lev[c->L - 1 - (int)level], index_dev, s));
    }
    return RC_OK;
}

static bool tiling_legal(int level, int shape, int rpt)
{
    return level >= 1 && shape >= 0 && shape < TILE_COUNT && tile_kd(shape) <= num_dirs(level) &&
           (rpt == 1 || rpt == 2 || rpt == 4);
}

rc_status rc_set_level_tiling(rc_ctx *c, uint32_t level, uint32_t shape, uint32_t rpt)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_set_level_tiling: ctx is NULL");
    if ((int)level >= c->L || !tiling_legal((int)level, (int)shape, (int)rpt))
        return fail(c, RC_ERR_INVALID, "rc_set_level_tiling: illegal level/shape/rays_per_lane");
    c->tile[level] = (int)shape;
    c->rpt[level] = (int)rpt;
    drop_graph(c);
    return RC_OK;
}

rc_status rc_set_jfa_fuse(rc_ctx *c, uint32_t fused_step)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_set_jfa_fuse: ctx is NULL");
    if (fused_step != 0 && fused_step != 1 && fused_step != 2 && fused_step != 4 && fused_step != 8)
        return fail(c, RC_ERR_INVALID, "rc_set_jfa_fuse: fused_step must be 0, 1, 2, 4 or 8");
    c->jfa_fuse = (int)fused_step;
    drop_graph(c);
    return RC_OK;
}

uint32_t rc_get_jfa_fuse(const rc_ctx *c) { return c ? (uint32_t)c->jfa_fuse : 0u; }

rc_status rc_set_jfa_rows(rc_ctx *c, uint32_t rows)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_set_jfa_rows: ctx is NULL");
    if (rows != 1 && rows != 2 && rows != 4) return fail(c, RC_ERR_INVALID, "rc_set_jfa_rows: rows must be 1, 2 or 4");
    c->jfa_rows = (int)rows;
    drop_graph(c);
    return RC_OK;
}

uint32_t rc_get_jfa_rows(const rc_ctx *c) { return c ? (uint32_t)c->jfa_rows : 0u; }

rc_status rc_set_sdf_mode(rc_ctx *c, uint32_t mode)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_set_sdf_mode: ctx is NULL");
    if (mode != RC_SDF_EXACT && mode != RC_SDF_JUMP_FLOOD)
        return fail(c, RC_ERR_INVALID, "rc_set_sdf_mode: mode must be RC_SDF_EXACT or RC_SDF_JUMP_FLOOD");
    if ((int)mode == c->sdf_mode) return RC_OK;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    const int old = c->sdf_mode;
    c->sdf_mode = (int)mode;
    const rc_status st = ensure_sdf_buffers(c);
    if (st != RC_OK) {
        c->sdf_mode = old;
        return st;
    }
    c->ran = false;  // the debug readbacks follow the next frame
    drop_graph(c);
    return RC_OK;
}

uint32_t rc_get_sdf_mode(const rc_ctx *c) { return c ? (uint32_t)c->sdf_mode : 0u; }

rc_status rc_set_march(rc_ctx *c, uint32_t split, uint32_t speculation, uint32_t rays_per_lane)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_set_march: ctx is NULL");
    if (split > 2 || (speculation != 1 && speculation != 2 && speculation != 4) ||
        (rays_per_lane != 1 && rays_per_lane != 2 && rays_per_lane != 4 && rays_per_lane != 8) ||
        (split == 2 && rays_per_lane != 1))
        return fail(c, RC_ERR_INVALID,
                    "rc_set_march: split 0/1/2, speculation 1/2/4, rays_per_lane 1/2/4/8 (1 with split 2)");
    c->march_split = (int)split;
    c->spec = (int)speculation;
    c->march_m = (int)rays_per_lane;
    drop_graph(c);
    return RC_OK;
}

rc_status rc_set_march_order(rc_ctx *c, uint32_t level, uint32_t order)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_set_march_order: ctx is NULL");
    if ((int)level >= c->L || !march_map_legal((int)level, (int)order))
        return fail(c, RC_ERR_INVALID, "rc_set_march_order: level 1 .. L-1, order 0 .. 5 with its slice at most N_level");
    c->mmap[level] = (int)order;
    drop_graph(c);
    return RC_OK;
}

rc_status rc_get_march_order(const rc_ctx *c, uint32_t level, uint32_t *order)
{
    if (!c || !order) return fail(nullptr, RC_ERR_NULL, "rc_get_march_order: NULL argument");
    if (level == 0 || (int)level >= c->L) return fail(nullptr, RC_ERR_INVALID, "rc_get_march_order: level out of range");
    *order = (uint32_t)c->mmap[level];
    return RC_OK;
}

rc_status rc_get_march(const rc_ctx *c, uint32_t *split, uint32_t *speculation, uint32_t *rays_per_lane)
{
    if (!c || !split || !speculation || !rays_per_lane) return fail(nullptr, RC_ERR_NULL, "rc_get_march: NULL argument");
    *split = (uint32_t)c->march_split;
    *speculation = (uint32_t)c->spec;
    *rays_per_lane = (uint32_t)c->march_m;
    return RC_OK;
}

rc_status rc_get_level_tiling(const rc_ctx *c, uint32_t level, uint32_t *shape, uint32_t *rpt)
{
    if (!c || !shape || !rpt) return fail(nullptr, RC_ERR_NULL, "rc_get_level_tiling: NULL argument");
    if (level == 0 || (int)level >= c->L) return fail(nullptr, RC_ERR_INVALID, "rc_get_level_tiling: level out of range");
    *shape = (uint32_t)c->tile[level];
    *rpt = (uint32_t)c->rpt[level];
    return RC_OK;
}

// Time every legal tiling of every level i >= 1 on this G-buffer and keep the
// fastest.  A level kernel's cost does not depend on the upper level's
// values, so after one full frame each candidate is launched in isolation.
rc_status rc_autotune(rc_ctx *c, const rc_gbuffer *g, float *irr, uint32_t iters, void *stream)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_autotune: ctx is NULL");
    if (!g || !g->depth || !g->emissive || !irr) return fail(c, RC_ERR_NULL, "rc_autotune: G-buffer or output is NULL");
    if (iters == 0) iters = 1;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    hipStream_t s = (hipStream_t)stream;
    const uint32_t saved_flags = c->cfg.flags;
    c->cfg.flags &= ~RC_FLAG_PROFILE;
    rc_status st = run_frame(c, g->depth, g->emissive, irr, s);
    c->cfg.flags = saved_flags;
    if (st != RC_OK) return st;
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    if (c->sdf_mode == RC_SDF_JUMP_FLOOD) {
        float best = 1e30f;
        int best_fuse = c->jfa_fuse, best_rows = c->jfa_rows;
        const int saved = c->jfa_fuse, saved_rows = c->jfa_rows;
        for (int rows : {1, 2, 4})
        for (int fuse : {0, 2, 4, 8}) {
            c->jfa_fuse = fuse;
            c->jfa_rows = rows;
            HIP_TRY(c, hipEventRecord(e0, s));
            for (uint32_t it = 0; it < iters; ++it) {
                rc_status r = run_frame(c, g->depth, g->emissive, irr, s, true);
                if (r != RC_OK) { c->jfa_fuse = saved; c->jfa_rows = saved_rows; return r; }
            }
            HIP_TRY(c, hipEventRecord(e1, s));
            HIP_TRY(c, hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; best_fuse = fuse; best_rows = rows; }
        }
        c->jfa_fuse = best_fuse;
        c->jfa_rows = best_rows;
    }
    // A frame with one level marches nothing beyond level 0: there is no
    // march mode to tune (and no split march to run).
    if (c->L > 1) {
        float best = 1e30f;
        int best_split = c->march_split, best_spec = c->spec, best_m = c->march_m;
        for (int split : {0, 1, 2})
        for (int spec : {1, 2, 4})
            for (int m : {1, 2, 4, 8}) {
                if (!split && (spec != 1 || m != 1)) continue;
                if (split == 2 && m != 1) continue;
                c->march_split = split;
                c->spec = spec;
                c->march_m = m;
                HIP_TRY(c, hipEventRecord(e0, s));
                for (uint32_t it = 0; it < iters; ++it) {
                    rc_status r = run_frame(c, g->depth, g->emissive, irr, s, true);
                    if (r != RC_OK) return r;
                }
                HIP_TRY(c, hipEventRecord(e1, s));
                HIP_TRY(c, hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; best_split = split; best_spec = spec; best_m = m; }
            }
        c->march_split = best_split;
        c->spec = best_spec;
        c->march_m = best_m;
        if (best_split == 1) {
            // the order of each level's rays over the lanes, level by level
            // (the march launch timed whole, the other levels as chosen)
            for (int i = c->L - 1; i >= 1; --i) {
                float lbest = 1e30f;
                int bmap = c->mmap[i];
                for (int map = 0; map < kMarchMaps; ++map) {
                    if (!march_map_legal(i, map)) continue;
                    c->mmap[i] = map;
                    const MarchAllArgs m = march_all_args(c);
                    HIP_TRY(c, launch_march_all(m, c->spec, c->march_m, s));  // warm
                    HIP_TRY(c, hipEventRecord(e0, s));
                    for (uint32_t it = 0; it < iters; ++it) HIP_TRY(c, launch_march_all(m, c->spec, c->march_m, s));
                    HIP_TRY(c, hipEventRecord(e1, s));
                    HIP_TRY(c, hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
                    if (ms < lbest) { lbest = ms; bmap = map; }
                }
                c->mmap[i] = bmap;
            }
        }
        if (best_split == 2) HIP_TRY(c, launch_march_persist(march_all_args(c), c->spec, s));
        else if (best_split) HIP_TRY(c, launch_march_all(march_all_args(c), c->spec, c->march_m, s));
    }
    for (int i = 1; i < c->L; ++i) {
        float best = 1e30f;
        int best_shape = c->tile[i], best_rpt = c->rpt[i];
        LevelArgs a = level_args(c, i, g->emissive);
        a.emis4 = c->emis4;  // as the frame runs them (filled by the frames above)
        if (c->march_split) a.codes = c->codes + c->code_off[i];
        for (int shape = 0; shape < TILE_COUNT; ++shape) {
            for (int rpt : {1, 2, 4}) {
                if (!tiling_legal(i, shape, rpt)) continue;
                HIP_TRY(c, launch_level(a, c->storage, shape, rpt, s));  // warm
                HIP_TRY(c, hipEventRecord(e0, s));
                for (uint32_t it = 0; it < iters; ++it) HIP_TRY(c, launch_level(a, c->storage, shape, rpt, s));
                HIP_TRY(c, hipEventRecord(e1, s));
                HIP_TRY(c, hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; best_shape = shape; best_rpt = rpt; }
            }
        }
        c->tile[i] = best_shape;
        c->rpt[i] = best_rpt;
    }
    if (c->surr_mode && c->mlp_ready) {
        // the learned frame's level 1 is another kernel variant (marched in
        // place, no ray codes): tune it on a learned frame's level 2
        st = ensure_surrogate(c);
        if (st != RC_OK) return st;
        st = run_frame_surrogate(c, g->depth, g->emissive, irr, s, true);
        if (st != RC_OK) return st;
        LevelArgs a1 = level_args(c, 1, g->emissive);
        a1.sdf = c->sdf_near;
        a1.upper = c->surr_a2;
        a1.out = c->surr_a1;
        a1.codes = nullptr;
        float best = 1e30f;
        int best_shape = c->tile[1], best_rpt = c->rpt[1];
        for (int shape = 0; shape < TILE_COUNT; ++shape) {
            for (int rpt : {1, 2, 4}) {
                if (!tiling_legal(1, shape, rpt)) continue;
                HIP_TRY(c, launch_level(a1, c->storage, shape, rpt, s));  // warm
                HIP_TRY(c, hipEventRecord(e0, s));
                for (uint32_t it = 0; it < iters; ++it) HIP_TRY(c, launch_level(a1, c->storage, shape, rpt, s));
                HIP_TRY(c, hipEventRecord(e1, s));
                HIP_TRY(c, hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; best_shape = shape; best_rpt = rpt; }
            }
        }
        c->surr_tile1 = best_shape;
        c->surr_rpt1 = best_rpt;
    }
    drop_graph(c);
    // leave the level buffers consistent with one full frame
    return run_frame(c, g->depth, g->emissive, irr, s);
}

rc_status rc_error_norms(rc_ctx *c, const float *x, const void *ref, uint32_t ref_is_f64, uint64_t n,
                         double *out_dev, void *stream)
{
    if (!c || !x || !ref || !out_dev) return fail(c, RC_ERR_NULL, "rc_error_norms: NULL argument");
    if (n == 0) return fail(c, RC_ERR_INVALID, "rc_error_norms: empty input");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, launch_error_norms(x, ref, ref_is_f64 != 0, (size_t)n, c->norm_ws, out_dev, (hipStream_t)stream));
    return RC_OK;
}

void rc_srgb_thresholds(float *t)
{
    if (t) srgb_thresholds(t);
}

rc_status rc_composite(rc_ctx *c, const float *albedo, const float *emissive, const float *irradiance,
                       const float *direct, uint32_t *bgra8, void *stream)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_composite: ctx is NULL");
    if (!albedo || !emissive || !irradiance || !bgra8) return fail(c, RC_ERR_NULL, "rc_composite: NULL buffer");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (!c->srgb_table) {
        float t[256];
        srgb_thresholds(t);
        HIP_TRY(c, hipMalloc(&c->srgb_table, sizeof(t)));
        HIP_TRY(c, hipMemcpy(c->srgb_table, t, sizeof(t), hipMemcpyHostToDevice));
    }
    HIP_TRY(c, launch_composite(albedo, emissive, irradiance, direct, c->srgb_table, bgra8, (size_t)c->W * c->H,
                                (hipStream_t)stream));
    return RC_OK;
}

rc_status rc_probe_grid(uint32_t W, uint32_t H, uint32_t level, uint32_t *px, uint32_t *py)
{
    if (!px || !py) return fail(nullptr, RC_ERR_NULL, "rc_probe_grid: NULL argument");
    if (!dims_ok(W, H) || level >= RC_MAX_LEVELS) return fail(nullptr, RC_ERR_INVALID, "rc_probe_grid: out of range");
    *px = (uint32_t)grid_dim((int)W, (int)level);
    *py = (uint32_t)grid_dim((int)H, (int)level);
    return RC_OK;
}

uint64_t rc_level_rays(uint32_t W, uint32_t H, uint32_t level)
{
    if (!dims_ok(W, H) || level >= RC_MAX_LEVELS) return 0;
    return (uint64_t)grid_dim((int)W, (int)level) * grid_dim((int)H, (int)level) * num_dirs((int)level);
}

uint32_t rc_levels_for_coverage(uint32_t W, uint32_t H) { return dims_ok(W, H) ? coverage_levels(W, H) : 0; }

rc_status rc_scene_generate(uint64_t seed, uint32_t n_boxes, uint32_t n_emissive, rc_box *boxes)
{
    if (n_boxes > 0 && !boxes) return fail(nullptr, RC_ERR_NULL, "rc_scene_generate: boxes is NULL");
    if (n_boxes > RC_MAX_BOXES || n_emissive > n_boxes) return fail(nullptr, RC_ERR_INVALID, "rc_scene_generate: bad counts");
    scene_generate(seed, (int)n_boxes, (int)n_emissive, boxes);
    return RC_OK;
}

rc_status rc_gbuffer_raster(const rc_box *boxes, uint32_t n, uint32_t W, uint32_t H, float *depth, float *albedo,
                            float *emissive, void *stream)
{
    if ((n > 0 && !boxes) || !depth || !emissive) return fail(nullptr, RC_ERR_NULL, "rc_gbuffer_raster: NULL argument");
    if (!dims_ok(W, H) || n > RC_MAX_BOXES) return fail(nullptr, RC_ERR_INVALID, "rc_gbuffer_raster: out of range");
    int ndev = 0;
    if (rc_device_count(&ndev) != RC_OK) return RC_ERR_NO_DEVICE;
    double *dev = nullptr;
    const size_t bytes = (size_t)(n ? n : 1) * sizeof(rc_box);
    hipError_t e = hipMallocAsync((void **)&dev, bytes, (hipStream_t)stream);
    if (e != hipSuccess) return fail(nullptr, RC_ERR_HIP, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    if (n) e = hipMemcpyAsync(dev, boxes, n * sizeof(rc_box), hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e == hipSuccess) e = launch_raster(dev, (int)n, (int)W, (int)H, depth, albedo, emissive, (hipStream_t)stream);
    hipError_t e2 = hipFreeAsync(dev, (hipStream_t)stream);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return fail(nullptr, RC_ERR_HIP, std::string("rc_gbuffer_raster: ") + hipGetErrorString(e));
    return RC_OK;
}

// ---- learned far field --------------------------------------------------------
rc_status rc_set_surrogate(rc_ctx *c, const rc_surrogate_config *cfg)
{
    if (!c || !cfg) return fail(c, RC_ERR_NULL, "rc_set_surrogate: NULL argument");
    if (cfg->version != RC_ABI_VERSION || cfg->mode > 1 || cfg->threshold != cfg->threshold)
        return fail(c, RC_ERR_INVALID, "rc_set_surrogate: bad version, mode or threshold");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (cfg->mode == 1) {
        rc_status st = ensure_surrogate(c);
        if (st != RC_OK) return st;
    }
    c->surr_mode = (int)cfg->mode;
    c->surr_thr = cfg->threshold;
    drop_graph(c);
    return RC_OK;
}

rc_status rc_get_surrogate(const rc_ctx *c, rc_surrogate_config *out)
{
    if (!c || !out) return fail(nullptr, RC_ERR_NULL, "rc_get_surrogate: NULL argument");
    out->version = RC_ABI_VERSION;
    out->mode = (uint32_t)c->surr_mode;
    out->threshold = c->surr_thr;
    out->reserved = 0;
    return RC_OK;
}

rc_status rc_mlp_set(rc_ctx *c, const float *params, const float *mean, const float *istd, float out_scale)
{
    if (!c || !params || !mean || !istd) return fail(c, RC_ERR_NULL, "rc_mlp_set: NULL argument");
    if (!(out_scale > 0.0f) || !(out_scale < __builtin_huge_valf()))
        return fail(c, RC_ERR_INVALID, "rc_mlp_set: out_scale must be positive and finite");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    return mlp_upload(c, params, mean, istd, out_scale);
}

rc_status rc_mlp_get(const rc_ctx *c, float *params, float *mean, float *istd, float *out_scale)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_mlp_get: ctx is NULL");
    rc_ctx *m = const_cast<rc_ctx *>(c);
    HIP_TRY(m, hipSetDevice(c->cfg.device));
    HIP_TRY(m, hipDeviceSynchronize());
    if (params) HIP_TRY(m, hipMemcpy(params, c->mlp.master, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    if (mean) HIP_TRY(m, hipMemcpy(mean, c->mlp.in_mean, kMlpIn * sizeof(float), hipMemcpyDeviceToHost));
    if (istd) HIP_TRY(m, hipMemcpy(istd, c->mlp.in_istd, kMlpIn * sizeof(float), hipMemcpyDeviceToHost));
    if (out_scale) *out_scale = c->out_scale;
    return RC_OK;
}

// blob: magic[8], u32 version, inputs, outputs, params, f32 out_scale, u32
// train_scenes, f32 in_mean[64], in_istd[64], params[16640]
static const uint32_t kBlobVersion = 2;  // 2: 64 inputs (cone transmittance added)
static size_t blob_bytes() { return 8 + 6 * 4 + (2 * kMlpIn + kMlpParams) * sizeof(float); }

rc_status rc_surrogate_save(const rc_ctx *c, void *buf, uint64_t cap, uint64_t *size)
{
    if (!c || !size) return fail(nullptr, RC_ERR_NULL, "rc_surrogate_save: NULL argument");
    *size = blob_bytes();
    if (!buf) return RC_OK;
    rc_ctx *m = const_cast<rc_ctx *>(c);
    if (cap < blob_bytes()) return fail(m, RC_ERR_INVALID, "rc_surrogate_save: buffer too small");
    if (!c->mlp_ready) return fail(m, RC_ERR_STATE, "rc_surrogate_save: no network");
    std::vector<float> params(kMlpParams), mean(kMlpIn), istd(kMlpIn);
    float os = 0.f;
    rc_status st = rc_mlp_get(c, params.data(), mean.data(), istd.data(), &os);
    if (st != RC_OK) return st;
    uint8_t *p = (uint8_t *)buf;
    memcpy(p, RC_SURROGATE_BLOB_MAGIC, 8);
    const uint32_t hdr[4] = {kBlobVersion, (uint32_t)kMlpIn, (uint32_t)kMlpOut, (uint32_t)kMlpParams};
    memcpy(p + 8, hdr, 16);
    memcpy(p + 24, &os, 4);
    memcpy(p + 28, &c->train_scenes, 4);
    memcpy(p + 32, mean.data(), kMlpIn * 4);
    memcpy(p + 32 + kMlpIn * 4, istd.data(), kMlpIn * 4);
    memcpy(p + 32 + 2 * kMlpIn * 4, params.data(), kMlpParams * 4);
    return RC_OK;
}

rc_status rc_surrogate_load(rc_ctx *c, const void *buf, uint64_t size)
{
    if (!c || !buf) return fail(c, RC_ERR_NULL, "rc_surrogate_load: NULL argument");
    const uint8_t *p = (const uint8_t *)buf;
    if (size != blob_bytes() || memcmp(p, RC_SURROGATE_BLOB_MAGIC, 8) != 0)
        return fail(c, RC_ERR_INVALID, "rc_surrogate_load: not a weights blob of this format");
    uint32_t hdr[4];
    memcpy(hdr, p + 8, 16);
    if (hdr[0] != kBlobVersion || hdr[1] != (uint32_t)kMlpIn || hdr[2] != (uint32_t)kMlpOut ||
        hdr[3] != (uint32_t)kMlpParams)
        return fail(c, RC_ERR_INVALID, "rc_surrogate_load: version or network shape mismatch");
    float os;
    memcpy(&os, p + 24, 4);
    std::vector<float> params(kMlpParams), mean(kMlpIn), istd(kMlpIn);
    memcpy(mean.data(), p + 32, kMlpIn * 4);
    memcpy(istd.data(), p + 32 + kMlpIn * 4, kMlpIn * 4);
    memcpy(params.data(), p + 32 + 2 * kMlpIn * 4, kMlpParams * 4);
    rc_status st = rc_mlp_set(c, params.data(), mean.data(), istd.data(), os);
    if (st == RC_OK) memcpy(&c->train_scenes, p + 28, 4);
    return st;
}

rc_status rc_mlp_forward(rc_ctx *c, const float *feats, uint32_t n, float *out, void *stream)
{
    if (!c || !feats || !out) return fail(c, RC_ERR_NULL, "rc_mlp_forward: NULL argument");
    if (!c->mlp_ready) return fail(c, RC_ERR_STATE, "rc_mlp_forward: no network");
    if (n > (1u << 30)) return fail(c, RC_ERR_INVALID, "rc_mlp_forward: n too large");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, launch_mlp_infer(c->mlp, feats, (int)n, out, (hipStream_t)stream));
    return RC_OK;
}

rc_status rc_mlp_gradient(rc_ctx *c, const float *feats, const float *targets, uint32_t n, float *grad, float *loss,
                          void *stream)
{
    if (!c || !feats || !targets || !grad || !loss) return fail(c, RC_ERR_NULL, "rc_mlp_gradient: NULL argument");
    if (!c->mlp_ready) return fail(c, RC_ERR_STATE, "rc_mlp_gradient: no network");
    if (n < 1 || n > (1u << 22)) return fail(c, RC_ERR_INVALID, "rc_mlp_gradient: n out of range (1 .. 2^22)");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    hipStream_t s = (hipStream_t)stream;
    const int parts = (int)((n + 31) / 32);
    rc_status st = ensure_mlp_scratch(c, parts, (int)n);
    if (st != RC_OK) return st;
    if ((int)n > 0) {
        if (c->mlp_idx) hipFree(c->mlp_idx);
        c->mlp_idx = nullptr;
        std::vector<int> idx(n);
        for (uint32_t i = 0; i < n; ++i) idx[i] = (int)i;
        HIP_TRY(c, hipMalloc(&c->mlp_idx, n * sizeof(int)));
        HIP_TRY(c, hipMemcpy(c->mlp_idx, idx.data(), n * sizeof(int), hipMemcpyHostToDevice));
    }
    HIP_TRY(c, launch_mlp_train(c->mlp, feats, targets, c->mlp_idx, (int)n, 1, c->mlp_partial, c->mlp_loss_part, parts,
                                s));
    HIP_TRY(c, launch_mlp_reduce(c->mlp_partial, parts, grad, s));
    HIP_TRY(c, launch_sum_f32(c->mlp_loss_part, parts, loss, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return RC_OK;
}

static rc_status adam_step(rc_ctx *c, const float *grad, float lr, hipStream_t s)
{
    const double b1 = 0.9, b2 = 0.999;
    c->adam_t += 1;
    const float c1 = (float)(1.0 - pow(b1, c->adam_t)), c2 = (float)(1.0 - pow(b2, c->adam_t));
    HIP_TRY(c, launch_mlp_adam(c->mlp, grad, lr, (float)b1, (float)b2, 1e-8f, c1, c2, s));
    return RC_OK;
}

rc_status rc_mlp_adam(rc_ctx *c, const float *grad, float lr, void *stream)
{
    if (!c || !grad) return fail(c, RC_ERR_NULL, "rc_mlp_adam: NULL argument");
    if (!c->mlp_ready) return fail(c, RC_ERR_STATE, "rc_mlp_adam: no network");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    rc_status st = adam_step(c, grad, lr, (hipStream_t)stream);
    if (st == RC_OK) drop_graph(c);
    return st;
}

rc_status rc_surrogate_samples(rc_ctx *c, const rc_gbuffer *g, float *feats, float *targets, void *stream)
{
    if (!c || !g || !g->depth || !g->emissive || !feats || !targets)
        return fail(c, RC_ERR_NULL, "rc_surrogate_samples: NULL argument");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    return gen_samples(c, g->depth, g->emissive, feats, targets, (hipStream_t)stream);
}

// He-uniform initial weights (limit sqrt(6 / fan_in)) from SplitMix64; the
// output layer's radiance columns are halved and its padding columns zero;
// biases zero.
static void init_params(uint64_t seed, std::vector<float> &p)
{
    uint64_t st = seed;
    p.assign(kMlpParams, 0.0f);
    size_t off = 0;
    for (int l = 0; l < 4; ++l) {
        const int K = l == 0 ? kMlpIn : kMlpHid;
        const double lim = sqrt(6.0 / K);
        for (int i = 0; i < K; ++i)
            for (int o = 0; o < 64; ++o) {
                const double u = (double)(mix64(st) >> 11) * 0x1.0p-53;
                double w = (2.0 * u - 1.0) * lim;
                if (l == 3 && o > kMlpErr) w = 0.0;
                if (l == 3 && o < kMlpRad) w *= 0.5;
                p[off + (size_t)i * 64 + o] = (float)w;
            }
        off += (size_t)K * 64 + 64;
    }
}

static rc_status train_rows(rc_ctx *c, const float *feats, float *targets, size_t N, const rc_train_config *cfg,
                            rc_train_report *rep, std::chrono::steady_clock::time_point t_start, uint32_t scenes);

// the optimiser fields shared by both trainers
static bool train_cfg_ok(const rc_train_config *cfg)
{
    return cfg->version == RC_ABI_VERSION && cfg->epochs >= 1 && cfg->epochs <= RC_TRAIN_MAX_EPOCHS &&
           cfg->batch >= 32 && cfg->batch % 32 == 0 && cfg->batch <= (1u << 22) && cfg->lr > 0.0f;
}

rc_status rc_surrogate_train(rc_ctx *c, const rc_train_config *cfg, rc_train_report *rep)
{
    if (!c || !cfg) return fail(c, RC_ERR_NULL, "rc_surrogate_train: NULL argument");
    if (!train_cfg_ok(cfg) || cfg->scenes < 1 || cfg->levels < 3 || cfg->levels > RC_MAX_LEVELS ||
        !dims_ok(cfg->width, cfg->height) || cfg->n_boxes > RC_MAX_BOXES || cfg->n_emissive > cfg->n_boxes)
        return fail(c, RC_ERR_INVALID, "rc_surrogate_train: bad configuration");
    const size_t rows = (size_t)grid_dim(cfg->width, 2) * grid_dim(cfg->height, 2) * cfg->scenes;
    if (rows > (size_t)INT32_MAX) return fail(c, RC_ERR_INVALID, "rc_surrogate_train: more than 2^31 - 1 rows");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    auto now = []() { return std::chrono::steady_clock::now(); };
    const auto t_start = now();

    // data: a plain context at the training size generates the rows
    rc_config tcfg{RC_ABI_VERSION, cfg->width, cfg->height, cfg->levels, RC_STORAGE_F32, 0,
                   c->cfg.device, 0};
    rc_ctx *tc = nullptr;
    rc_status st = rc_create(&tcfg, &tc);
    if (st != RC_OK) return fail(c, st, std::string("rc_surrogate_train: ") + rc_last_error(nullptr));
    struct Guard {
        rc_ctx *tc;
        std::vector<void *> dev;
        ~Guard()
        {
            for (void *p : dev)
                if (p) hipFree(p);
            rc_destroy(tc);
        }
    } guard{tc, {}};
    const size_t px = (size_t)cfg->width * cfg->height;
    const size_t P = (size_t)grid_dim(cfg->width, 2) * grid_dim(cfg->height, 2);  // rows per scene
    const size_t N = P * cfg->scenes;
    float *depth = nullptr, *emis = nullptr, *feats = nullptr, *targets = nullptr;
    double *boxes = nullptr;
    HIP_TRY(c, hipMalloc(&depth, px * sizeof(float)));
    guard.dev.push_back(depth);
    HIP_TRY(c, hipMalloc(&emis, 3 * px * sizeof(float)));
    guard.dev.push_back(emis);
    HIP_TRY(c, hipMalloc(&feats, N * kMlpIn * sizeof(float)));
    guard.dev.push_back(feats);
    HIP_TRY(c, hipMalloc(&targets, N * kMlpRad * sizeof(float)));
    guard.dev.push_back(targets);
    const size_t nb_scene = cfg->n_boxes ? cfg->n_boxes : 1;
    std::vector<rc_box> hb((size_t)cfg->scenes * nb_scene);
    for (uint32_t sc = 0; sc < cfg->scenes; ++sc)
        scene_generate(cfg->scene_seed0 + sc, (int)cfg->n_boxes, (int)cfg->n_emissive, hb.data() + sc * nb_scene);
    HIP_TRY(c, hipMalloc(&boxes, hb.size() * sizeof(rc_box)));
    guard.dev.push_back(boxes);
    HIP_TRY(c, hipMemcpy(boxes, hb.data(), hb.size() * sizeof(rc_box), hipMemcpyHostToDevice));
    hipStream_t s = nullptr;
    for (uint32_t sc = 0; sc < cfg->scenes; ++sc) {
        HIP_TRY(c, launch_raster(boxes + sc * nb_scene * (sizeof(rc_box) / sizeof(double)), (int)cfg->n_boxes,
                                 (int)cfg->width, (int)cfg->height, depth, nullptr, emis, s));
        st = gen_samples(tc, depth, emis, feats + sc * P * kMlpIn, targets + sc * P * kMlpRad, s);
        if (st != RC_OK) return fail(c, st, std::string("rc_surrogate_train: ") + tc->err);
    }
    HIP_TRY(c, hipDeviceSynchronize());
    return train_rows(c, feats, targets, N, cfg, rep, t_start, cfg->scenes);
}

// Normalisation, initial weights and the epochs of Adam steps on N rows
// already on the device (targets are scaled in place by 1 / out_scale).
static rc_status train_rows(rc_ctx *c, const float *feats, float *targets, size_t N, const rc_train_config *cfg,
                            rc_train_report *rep, std::chrono::steady_clock::time_point t_start, uint32_t scenes)
{
    auto now = []() { return std::chrono::steady_clock::now(); };
    std::vector<void *> owned;
    struct Free {
        std::vector<void *> &v;
        ~Free()
        {
            for (void *p : v)
                if (p) hipFree(p);
        }
    } free_guard{owned};
    hipStream_t s = nullptr;
    double *stats = nullptr;
    int *idx = nullptr;
    float *partial = nullptr, *loss_part = nullptr, *grad = nullptr, *step_loss = nullptr;
    rc_status st;
    // normalisation: per-input mean and std, one RMS scale for the targets
    const int parts = 1024;
    HIP_TRY(c, hipMalloc(&stats, (size_t)parts * 2 * kMlpIn * sizeof(double)));
    owned.push_back(stats);
    std::vector<double> hs((size_t)parts * 2 * kMlpIn);
    std::vector<float> mean(kMlpIn), istd(kMlpIn);
    HIP_TRY(c, launch_col_stats(feats, N, kMlpIn, stats, parts, s));
    HIP_TRY(c, hipMemcpy(hs.data(), stats, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < kMlpIn; ++j) {
        double a = 0.0, b = 0.0;
        for (int p = 0; p < parts; ++p) {
            a += hs[(size_t)p * 2 * kMlpIn + j];
            b += hs[(size_t)p * 2 * kMlpIn + kMlpIn + j];
        }
        const double mu = a / (double)N, var = fmax(b / (double)N - mu * mu, 0.0);
        mean[j] = (float)mu;
        istd[j] = (float)(1.0 / fmax(sqrt(var), 1e-3));
    }
    HIP_TRY(c, launch_col_stats(targets, N, kMlpRad, stats, parts, s));
    HIP_TRY(c, hipMemcpy(hs.data(), stats, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
    double sq = 0.0;
    for (int p = 0; p < parts; ++p)
        for (int j = 0; j < kMlpRad; ++j) sq += hs[(size_t)p * 2 * kMlpRad + kMlpRad + j];
    const float out_scale = (float)fmax(sqrt(sq / ((double)N * kMlpRad)), 1e-6);
    HIP_TRY(c, launch_scale(targets, N * kMlpRad, 1.0f / out_scale, s));
    HIP_TRY(c, hipDeviceSynchronize());
    const auto t_gen = now();

    // network
    std::vector<float> params;
    init_params(cfg->seed, params);
    st = mlp_upload(c, params.data(), mean.data(), istd.data(), out_scale);
    if (st != RC_OK) return st;
    const int batch = (int)cfg->batch;
    const int blocks = batch / 32;
    const int nb = blocks > 1024 ? (blocks + 1023) / 1024 : 1;
    const int wgs = (blocks + nb - 1) / nb;
    const size_t steps = N / (size_t)batch;
    if (steps < 1) return fail(c, RC_ERR_INVALID, "rc_surrogate_train: fewer samples than one batch");
    HIP_TRY(c, hipMalloc(&idx, N * sizeof(int)));
    owned.push_back(idx);
    HIP_TRY(c, hipMalloc(&partial, (size_t)wgs * kMlpParams * sizeof(float)));
    owned.push_back(partial);
    HIP_TRY(c, hipMalloc(&loss_part, (size_t)wgs * sizeof(float)));
    owned.push_back(loss_part);
    HIP_TRY(c, hipMalloc(&grad, kMlpParams * sizeof(float)));
    owned.push_back(grad);
    HIP_TRY(c, hipMalloc(&step_loss, steps * sizeof(float)));
    owned.push_back(step_loss);
    std::vector<int> perm(N);
    std::vector<float> hl(steps);
    uint64_t rs = cfg->seed ^ 0x5DEECE66Dull;
    if (rep) {
        memset(rep, 0, sizeof(*rep));
        rep->version = RC_ABI_VERSION;
    }
    for (uint32_t e = 0; e < cfg->epochs; ++e) {
        for (size_t i = 0; i < N; ++i) perm[i] = (int)i;
        for (size_t i = N - 1; i > 0; --i) {
            const size_t j = (size_t)(mix64(rs) % (uint64_t)(i + 1));
            const int t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
        HIP_TRY(c, hipMemcpy(idx, perm.data(), N * sizeof(int), hipMemcpyHostToDevice));
        const float lr = cfg->lr * powf(cfg->lr_decay > 0.0f ? cfg->lr_decay : 1.0f, (float)e);
        for (size_t k = 0; k < steps; ++k) {
            HIP_TRY(c, launch_mlp_train(c->mlp, feats, targets, idx + k * batch, batch, nb, partial, loss_part, wgs, s));
            HIP_TRY(c, launch_mlp_reduce(partial, wgs, grad, s));
            HIP_TRY(c, launch_sum_f32(loss_part, wgs, step_loss + k, s));
            st = adam_step(c, grad, lr, s);
            if (st != RC_OK) return st;
        }
        HIP_TRY(c, hipMemcpy(hl.data(), step_loss, steps * sizeof(float), hipMemcpyDeviceToHost));
        double tot = 0.0;
        for (float v : hl) tot += v;
        if (rep) rep->epoch_loss[e] = (float)(tot / (double)steps);
    }
    HIP_TRY(c, hipDeviceSynchronize());
    c->train_scenes = scenes;
    c->mlp_ready = true;
    drop_graph(c);
    if (rep) {
        rep->epochs = cfg->epochs;
        rep->samples = N;
        rep->steps = steps * cfg->epochs;
        rep->gen_ms = std::chrono::duration<double, std::milli>(t_gen - t_start).count();
        rep->train_ms = std::chrono::duration<double, std::milli>(now() - t_gen).count();
        rep->out_scale = out_scale;
    }
    return RC_OK;
}

rc_status rc_surrogate_train_rows(rc_ctx *c, const float *feats, float *targets, uint64_t n_rows,
                                  const rc_train_config *cfg, rc_train_report *rep)
{
    if (!c || !cfg || !feats || !targets) return fail(c, RC_ERR_NULL, "rc_surrogate_train_rows: NULL argument");
    if (!train_cfg_ok(cfg)) return fail(c, RC_ERR_INVALID, "rc_surrogate_train_rows: bad configuration");
    if (n_rows < cfg->batch) return fail(c, RC_ERR_INVALID, "rc_surrogate_train_rows: fewer rows than one batch");
    if (n_rows > (uint64_t)INT32_MAX) return fail(c, RC_ERR_INVALID, "rc_surrogate_train_rows: more than 2^31 - 1 rows");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipDeviceSynchronize());
    return train_rows(c, feats, targets, (size_t)n_rows, cfg, rep, std::chrono::steady_clock::now(), cfg->scenes);
}

rc_status rc_debug_surrogate(rc_ctx *c, float *cone, float *level1, float *tile_err, void *stream)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_debug_surrogate: ctx is NULL");
    if (!c->surr_ran) return fail(c, RC_ERR_STATE, "rc_debug_surrogate: no learned-far-field frame yet");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    const size_t Q = (size_t)grid_dim(c->W, 2) * grid_dim(c->H, 2), P = (size_t)grid_dim(c->W, 1) * grid_dim(c->H, 1);
    // a frame keeps whole cones (near + far segment) only with the debug flag
    if (cone && !(c->cfg.flags & RC_FLAG_KEEP_DEBUG))
        return fail(c, RC_ERR_STATE, "rc_debug_surrogate: cones need RC_FLAG_KEEP_DEBUG");
    if (cone) HIP_TRY(c, hipMemcpyAsync(cone, c->cone, Q * 16 * sizeof(float4), hipMemcpyDeviceToDevice, s));
    if (level1) HIP_TRY(c, launch_level_to_f32(c->surr_a1, c->storage, P * 4, level1, s));
    if (tile_err)
        HIP_TRY(c, hipMemcpyAsync(tile_err, c->tile_err, (size_t)c->surr_tiles * sizeof(float), hipMemcpyDeviceToDevice, s));
    return RC_OK;
}

rc_status rc_debug_mip(rc_ctx *c, uint32_t level, float *emission, float *coverage, void *stream)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_debug_mip: ctx is NULL");
    if (!c->surr_alloc || !c->surr_ran) return fail(c, RC_ERR_STATE, "rc_debug_mip: no learned-far-field frame yet");
    if (level < 1 || (int)level > c->mip.levels) return fail(c, RC_ERR_INVALID, "rc_debug_mip: level out of range");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    const size_t n = (size_t)c->mip.w[level] * c->mip.h[level];
    if (n && (emission || coverage)) HIP_TRY(c, launch_mip_unpack(c->mip, (int)level, emission, coverage, s));
    return RC_OK;
}

// ---- batches of views --------------------------------------------------------
// Work-list mode runs the batch through rc_batch.hip's engine; streams mode
// keeps one context per view and captures their frames on several streams.
struct rc_batch {
    uint32_t mode = RC_BATCH_WORKLIST;
    int device = 0;
    std::vector<rc_view> views;
    std::vector<uint32_t> levels;
    WorkList *wl = nullptr;
    // streams mode
    std::vector<rc_ctx *> ctx;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    hipStream_t cap = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint32_t kernel_nodes = 0;
    std::string err;
};

static rc_status bfail(rc_batch *b, rc_status s, const std::string &msg)
{
    if (b) b->err = msg;
    g_err = msg;
    return s;
}

#define BHIP_TRY(b, expr)                                                                            \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return bfail((b), RC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void rc_batch_destroy(rc_batch *b)
{
    if (!b) return;
    hipSetDevice(b->device);
    hipDeviceSynchronize();
    worklist_destroy(b->wl);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    for (auto *c : b->ctx) rc_destroy(c);
    for (auto st : b->streams) hipStreamDestroy(st);
    for (auto e : b->joins) hipEventDestroy(e);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->cap) hipStreamDestroy(b->cap);
    delete b;
}

rc_status rc_batch_create(const rc_view *views, uint32_t n, const rc_batch_config *cfg, rc_batch **out)
{
    if (!views || !cfg || !out) return bfail(nullptr, RC_ERR_NULL, "rc_batch_create: views, cfg or out is NULL");
    *out = nullptr;
    if (n == 0 || n > 4096) return bfail(nullptr, RC_ERR_INVALID, "rc_batch_create: 1 .. 4096 views");
    if (cfg->storage > RC_STORAGE_BF16) return bfail(nullptr, RC_ERR_INVALID, "rc_batch_create: unknown storage");
    if (cfg->mode > RC_BATCH_STREAMS) return bfail(nullptr, RC_ERR_INVALID, "rc_batch_create: unknown mode");
    if (cfg->mode == RC_BATCH_STREAMS && (cfg->streams == 0 || cfg->streams > 16))
        return bfail(nullptr, RC_ERR_INVALID, "rc_batch_create: 1 .. 16 streams");
    for (uint32_t i = 0; i < n; ++i) {
        if (!views[i].depth || !views[i].emissive || !views[i].irradiance)
            return bfail(nullptr, RC_ERR_NULL, "rc_batch_create: a view has a NULL buffer");
        if (!dims_ok(views[i].width, views[i].height) || views[i].levels > RC_MAX_LEVELS)
            return bfail(nullptr, RC_ERR_INVALID, "rc_batch_create: view " + std::to_string(i) + " out of range");
    }
    int ndev = 0;
    if (rc_device_count(&ndev) != RC_OK) return RC_ERR_NO_DEVICE;
    if (cfg->device < 0 || cfg->device >= ndev)
        return bfail(nullptr, RC_ERR_INVALID, "rc_batch_create: device ordinal out of range");
    rc_batch *b = new (std::nothrow) rc_batch();
    if (!b) return bfail(nullptr, RC_ERR_OOM, "rc_batch_create: host allocation failed");
    b->mode = cfg->mode;
    b->device = cfg->device;
    b->views.assign(views, views + n);
    for (uint32_t i = 0; i < n; ++i)
        b->levels.push_back(views[i].levels ? views[i].levels : coverage_levels(views[i].width, views[i].height));
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) {
        rc_batch_destroy(b);
        return bfail(nullptr, RC_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    }
    if (b->mode == RC_BATCH_WORKLIST) {
        std::string msg;
        e = worklist_create(views, n, b->levels.data(), (Storage)cfg->storage, cfg->device, &b->wl, msg);
        if (e != hipSuccess) {
            rc_batch_destroy(b);
            return bfail(nullptr, e == hipErrorOutOfMemory ? RC_ERR_OOM : (e == hipErrorInvalidValue ? RC_ERR_INVALID : RC_ERR_HIP), msg);
        }
        b->kernel_nodes = (uint32_t)worklist_launches(b->wl);
        *out = b;
        return RC_OK;
    }
    for (uint32_t i = 0; i < n; ++i) {
        rc_config c{RC_ABI_VERSION, views[i].width, views[i].height, b->levels[i], cfg->storage, 0, cfg->device, 0};
        rc_ctx *ctx = nullptr;
        rc_status st = rc_create(&c, &ctx);
        if (st != RC_OK) {
            std::string msg = std::string("rc_batch_create: view ") + std::to_string(i) + ": " + rc_last_error(nullptr);
            rc_batch_destroy(b);
            return bfail(nullptr, st, msg);
        }
        b->ctx.push_back(ctx);
    }
    e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (uint32_t k = 0; k < cfg->streams && e == hipSuccess; ++k) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) b->streams.push_back(st);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->joins.push_back(ev);
    }
    if (e != hipSuccess) {
        rc_batch_destroy(b);
        return bfail(nullptr, RC_ERR_HIP, std::string("rc_batch_create: ") + hipGetErrorString(e));
    }
    *out = b;
    return RC_OK;
}

rc_status rc_batch_view_levels(const rc_batch *b, uint32_t i, uint32_t *levels)
{
    if (!b || !levels) return bfail(nullptr, RC_ERR_NULL, "rc_batch_view_levels: NULL argument");
    if (i >= b->views.size()) return bfail(const_cast<rc_batch *>(b), RC_ERR_INVALID, "rc_batch_view_levels: no such view");
    *levels = b->levels[i];
    return RC_OK;
}

rc_status rc_batch_info(const rc_batch *b, uint32_t *kernel_nodes, uint64_t *workspace_bytes)
{
    if (!b) return bfail(nullptr, RC_ERR_NULL, "rc_batch_info: batch is NULL");
    if (kernel_nodes) *kernel_nodes = b->kernel_nodes;
    if (workspace_bytes) *workspace_bytes = b->wl ? (uint64_t)worklist_workspace_bytes(b->wl) : 0;
    return RC_OK;
}

rc_status rc_batch_autotune(rc_batch *b, uint32_t iters, void *stream)
{
    if (!b) return bfail(nullptr, RC_ERR_NULL, "rc_batch_autotune: batch is NULL");
    BHIP_TRY(b, hipSetDevice(b->device));
    if (b->wl) {
        std::string msg;
        const hipError_t e = worklist_autotune(b->wl, (int)iters, (hipStream_t)stream, msg);
        if (e != hipSuccess) return bfail(b, RC_ERR_HIP, "rc_batch_autotune: " + msg);
        b->kernel_nodes = (uint32_t)worklist_launches(b->wl);
        return RC_OK;
    }
    std::vector<int> leader(b->ctx.size(), -1);
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        for (size_t j = 0; j < i; ++j)
            if (leader[j] == (int)j && b->ctx[j]->W == b->ctx[i]->W && b->ctx[j]->H == b->ctx[i]->H &&
                b->ctx[j]->L == b->ctx[i]->L) {
                leader[i] = (int)j;
                break;
            }
        if (leader[i] >= 0) continue;
        leader[i] = (int)i;
        rc_gbuffer g{b->views[i].depth, nullptr, b->views[i].emissive};
        rc_status st = rc_autotune(b->ctx[i], &g, b->views[i].irradiance, iters, stream);
        if (st != RC_OK) return bfail(b, st, std::string("rc_batch_autotune: ") + b->ctx[i]->err);
    }
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        const rc_ctx *l = b->ctx[leader[i]];
        rc_ctx *c = b->ctx[i];
        for (int k = 0; k < RC_MAX_LEVELS; ++k) {
            c->tile[k] = l->tile[k];
            c->rpt[k] = l->rpt[k];
        }
        c->jfa_fuse = l->jfa_fuse;
        c->march_split = l->march_split;
        c->spec = l->spec;
        c->march_m = l->march_m;
    }
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    return RC_OK;
}

static uint32_t count_kernel_nodes(hipGraph_t g)
{
    size_t n = 0;
    if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess || n == 0) return 0;
    std::vector<hipGraphNode_t> nodes(n);
    if (hipGraphGetNodes(g, nodes.data(), &n) != hipSuccess) return 0;
    uint32_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nodes[i], &t) == hipSuccess && t == hipGraphNodeTypeKernel) ++k;
    }
    return k;
}

// Streams mode: fork the capture stream into the worker streams, each view's
// frame on stream (view mod streams), join them back.
static rc_status batch_capture(rc_batch *b)
{
    BHIP_TRY(b, hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    rc_status st = RC_OK;
    hipError_t e = hipEventRecord(b->fork, b->cap);
    for (size_t k = 0; k < b->streams.size() && e == hipSuccess; ++k) e = hipStreamWaitEvent(b->streams[k], b->fork, 0);
    for (size_t i = 0; i < b->ctx.size() && e == hipSuccess && st == RC_OK; ++i) {
        const rc_view &v = b->views[i];
        st = run_frame(b->ctx[i], v.depth, v.emissive, v.irradiance, b->streams[i % b->streams.size()], true);
    }
    for (size_t k = 0; k < b->streams.size() && e == hipSuccess; ++k) {
        e = hipEventRecord(b->joins[k], b->streams[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->cap, b->joins[k], 0);
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->cap, &g);
    if (st != RC_OK || e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        if (st != RC_OK) return bfail(b, st, "rc_batch_run: capture of a view failed");
        return bfail(b, RC_ERR_HIP, std::string("rc_batch_run: capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->kernel_nodes = count_kernel_nodes(g);
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->gexec = nullptr;
        return bfail(b, RC_ERR_HIP, std::string("rc_batch_run: hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return RC_OK;
}

rc_status rc_batch_run(rc_batch *b, void *stream)
{
    if (!b) return bfail(nullptr, RC_ERR_NULL, "rc_batch_run: batch is NULL");
    BHIP_TRY(b, hipSetDevice(b->device));
    if (b->wl) {
        std::string msg;
        const hipError_t e = worklist_run(b->wl, (hipStream_t)stream, msg);
        if (e != hipSuccess) return bfail(b, RC_ERR_HIP, msg);
        return RC_OK;
    }
    if (!b->gexec) {
        rc_status st = batch_capture(b);
        if (st != RC_OK) return st;
    }
    BHIP_TRY(b, hipGraphLaunch(b->gexec, (hipStream_t)stream));
    for (auto *c : b->ctx) {
        c->stats.runs += 1;
        c->ran = true;
    }
    return RC_OK;
}

const char *rc_batch_last_error(const rc_batch *b) { return b ? b->err.c_str() : g_err.c_str(); }

}  // extern "C"
