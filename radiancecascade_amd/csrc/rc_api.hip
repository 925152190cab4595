// rc_api.hip -- the C ABI (include/rc_gi.h): context, frame orchestration,
// profiling, parity readbacks, geometry and scene helpers.
//
// A frame is 1 seed kernel + (P - 4) global jump-flood passes + 1 fused
// tail (short passes + SDF) + (L - 1) level kernels + 1 level-0/gather
// kernel, all stream ordered on the caller's stream.  The context owns the
// jump-flood ping-pong buffers, the SDF, the merged levels 1..L-1 and the
// per-level direction tables; it allocates them at create/resize time only.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "rc_internal.hpp"

using namespace rc;

struct rc_ctx {
    rc_config cfg{};
    int W = 0, H = 0, L = 0;
    Storage storage = Storage::F32;
    uint32_t *seed[2] = {nullptr, nullptr};
    float *sdf = nullptr;
    void *level[RC_MAX_LEVELS] = {};
    size_t level_texels[RC_MAX_LEVELS] = {};
    float *dbg_level0 = nullptr;
    float2 *dirs = nullptr;
    size_t dir_off[RC_MAX_LEVELS] = {};
    int tile[RC_MAX_LEVELS] = {};
    int rpt[RC_MAX_LEVELS] = {};
    int jfa_fuse = 0;  // largest jump-flood step fused into the LDS tail (0 = none)
    // split march: levels >= 1 marched by one k_march_all launch into `codes`,
    // then merged top-down; otherwise each level kernel marches its own rays
    int march_split = 1;
    int spec = 1;      // speculation depth of the split march (1, 2, 4)
    int march_m = 1;   // rays per lane, marched back to back (1, 2, 4, 8)
    int *codes = nullptr;
    size_t code_off[RC_MAX_LEVELS] = {};
    double *norm_ws = nullptr;  // error-norm partials
    int dbg_seed_buf = -1;
    bool ran = false;
    // host-staging path (rc_update_host / rc_render_host)
    float *g_depth = nullptr, *g_emissive = nullptr, *irr = nullptr;
    bool have_gbuffer = false;
    // whole-frame hipGraph for rc_run_device, keyed on the caller's buffers
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    const void *g_key[3] = {nullptr, nullptr, nullptr};
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

// Default tile per level: the top level has no upper neighbourhood to stage,
// so it takes whole 256-direction slices of one probe (no idle lanes at ragged
// grid edges); lower levels take 4x4 probes x 16 directions, whose staged
// upper block is exactly one texel per ray.
static int default_tile(int level, int L)
{
    const int N = num_dirs(level);
    if (level == L - 1 && N >= 256) return TILE_1x1x256;
    if (N >= 16) return TILE_4x4x16;
    return TILE_8x8x4;
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
    if (c->sdf) hipFree(c->sdf);
    c->sdf = nullptr;
    for (auto &p : c->level) { if (p) hipFree(p); p = nullptr; }
    if (c->dbg_level0) hipFree(c->dbg_level0);
    c->dbg_level0 = nullptr;
    if (c->dirs) hipFree(c->dirs);
    c->dirs = nullptr;
    if (c->codes) hipFree(c->codes);
    c->codes = nullptr;
    if (c->norm_ws) hipFree(c->norm_ws);
    c->norm_ws = nullptr;
    if (c->g_depth) hipFree(c->g_depth);
    if (c->g_emissive) hipFree(c->g_emissive);
    if (c->irr) hipFree(c->irr);
    c->g_depth = c->g_emissive = c->irr = nullptr;
    c->have_gbuffer = false;
    c->ran = false;
}

static rc_status allocate(rc_ctx *c, uint32_t W, uint32_t H, uint32_t L)
{
    release(c);
    c->W = (int)W;
    c->H = (int)H;
    c->L = (int)L;
    const size_t n = (size_t)W * H;
    HIP_TRY(c, hipMalloc(&c->seed[0], n * sizeof(uint32_t)));
    HIP_TRY(c, hipMalloc(&c->seed[1], n * sizeof(uint32_t)));
    HIP_TRY(c, hipMalloc(&c->sdf, sdf_elems((int)W, (int)H) * sizeof(float)));
    const size_t tb = texel_bytes(c->storage);
    for (int i = 1; i < (int)L; ++i) {
        const size_t texels = (size_t)grid_dim(W, i) * grid_dim(H, i) * (num_dirs(i) / 4);
        c->level_texels[i] = texels;
        HIP_TRY(c, hipMalloc(&c->level[i], texels * tb));
    }
    c->level_texels[0] = (size_t)grid_dim(W, 0) * grid_dim(H, 0);
    if (c->cfg.flags & RC_FLAG_KEEP_DEBUG) HIP_TRY(c, hipMalloc(&c->dbg_level0, c->level_texels[0] * 16));
    // direction tables: theta = 2 pi (k + 1/2) / N in double, rounded to fp32
    std::vector<float2> dirs;
    for (int i = 0; i < (int)L; ++i) {
        c->dir_off[i] = dirs.size();
        const int N = num_dirs(i);
        for (int k = 0; k < N; ++k) {
            const double th = 2.0 * M_PI * ((double)k + 0.5) / (double)N;
            dirs.push_back(make_float2((float)cos(th), (float)sin(th)));
        }
        c->tile[i] = default_tile(i, (int)L);
        c->rpt[i] = 2;
    }
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

static MarchAllArgs march_all_args(const rc_ctx *c)
{
    MarchAllArgs m{};
    m.sdf = c->sdf;
    m.W = c->W;
    m.H = c->H;
    int blocks = 0;
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
        l.block_begin = blocks;
        const int64_t rays = (int64_t)l.PX * l.PY * l.N, per = (int64_t)kBlock * c->march_m;
        l.blocks = (int)((rays + per - 1) / per);
        blocks += l.blocks;
    }
    return m;
}

struct StageTimer {
    rc_ctx *c;
    hipStream_t s;
    bool on;
    void begin(int stage)
    {
        if (on) { hipEventRecord(c->ev[2 * stage], s); c->ev_used[stage] = true; }
    }
    void end(int stage)
    {
        if (on) hipEventRecord(c->ev[2 * stage + 1], s);
    }
};

static rc_status run_frame(rc_ctx *c, const float *depth, const float *emissive, float *irr, hipStream_t s,
                           bool quiet = false)
{
    const bool prof = !quiet && (c->cfg.flags & RC_FLAG_PROFILE) != 0;
    const bool dbg = (c->cfg.flags & RC_FLAG_KEEP_DEBUG) != 0;
    StageTimer tm{c, s, prof};
    for (auto &u : c->ev_used) u = false;

    tm.begin(RC_STAGE_SEED);
    HIP_TRY(c, launch_seed_init(depth, emissive, c->seed[0], c->W, c->H, s));
    tm.end(RC_STAGE_SEED);

    // Jump flood: global passes down to the fused LDS tail (if any); without a
    // tail the last global pass writes the SDF.  A 1x1 image has no pass.
    const int P = jfa_num_passes(c->W, c->H);
    const int tail = jfa_tail_step(c->W, c->H, c->jfa_fuse);
    int cur = 0, global_passes = 0;
    c->dbg_seed_buf = -1;
    tm.begin(RC_STAGE_JFA);
    for (int step = P > 0 ? 1 << (P - 1) : 0; step >= 1 && step > tail; step >>= 1) {
        const bool final_pass = tail == 0 && step == 1;
        uint32_t *dst = final_pass ? (dbg ? c->seed[cur ^ 1] : nullptr) : c->seed[cur ^ 1];
        HIP_TRY(c, launch_jfa_pass(c->seed[cur], dst, c->sdf, c->W, c->H, step, final_pass, s));
        if (final_pass) c->dbg_seed_buf = dbg ? (cur ^ 1) : -1;
        else cur ^= 1;
        ++global_passes;
    }
    tm.end(RC_STAGE_JFA);
    const bool tail_kernel = tail > 0 || P == 0;
    if (tail_kernel) {
        tm.begin(RC_STAGE_JFA_TAIL);
        HIP_TRY(c, launch_jfa_tail(c->seed[cur], dbg ? c->seed[cur ^ 1] : nullptr, c->sdf, c->W, c->H, tail, s));
        tm.end(RC_STAGE_JFA_TAIL);
        c->dbg_seed_buf = dbg ? (cur ^ 1) : -1;
    }

    const bool split = c->march_split && c->L > 1;
    if (split) {
        tm.begin(RC_STAGE_MARCH);
        HIP_TRY(c, launch_march_all(march_all_args(c), c->spec, c->march_m, s));
        tm.end(RC_STAGE_MARCH);
    }
    for (int i = c->L - 1; i >= 1; --i) {
        tm.begin(RC_STAGE_LEVEL0 + i);
        LevelArgs a = level_args(c, i, emissive);
        if (split) a.codes = c->codes + c->code_off[i];
        HIP_TRY(c, launch_level(a, c->storage, c->tile[i], c->rpt[i], s));
        tm.end(RC_STAGE_LEVEL0 + i);
    }
    tm.begin(RC_STAGE_LEVEL0);
    HIP_TRY(c, launch_level0_gather(level_args(c, 0, emissive), c->storage, irr, dbg ? c->dbg_level0 : nullptr, s));
    tm.end(RC_STAGE_LEVEL0);

    if (prof) {
        HIP_TRY(c, hipStreamSynchronize(s));
        for (int st = 0; st < RC_NUM_STAGES; ++st) {
            if (!c->ev_used[st]) continue;
            float ms = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[2 * st], c->ev[2 * st + 1]));
            const uint32_t n = st == RC_STAGE_JFA ? (uint32_t)global_passes : 1u;
            if (n == 0) continue;
            c->stats.stage_ms[st] += ms;
            c->stats.stage_launches[st] += n;
        }
    }
    if (!quiet) {
        c->stats.runs += 1;
        c->ran = true;
    }
    return RC_OK;
}

// The frame's 15 launches (at 1080p) replayed as one graph: profiling and
// debug runs, which record events or extra outputs, launch directly.
static bool graph_enabled(const rc_ctx *c)
{
    return (c->cfg.flags & (RC_FLAG_PROFILE | RC_FLAG_KEEP_DEBUG | RC_FLAG_NO_GRAPH)) == 0;
}

static rc_status run_frame_graph(rc_ctx *c, const float *depth, const float *emissive, float *irr, hipStream_t s)
{
    if (!c->gexec || c->g_key[0] != depth || c->g_key[1] != emissive || c->g_key[2] != irr) {
        drop_graph(c);
        if (!c->cap_stream) HIP_TRY(c, hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
        HIP_TRY(c, hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
        rc_status st = run_frame(c, depth, emissive, irr, c->cap_stream, true);
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
    if (cfg->flags & ~(RC_FLAG_PROFILE | RC_FLAG_KEEP_DEBUG | RC_FLAG_NO_GRAPH)) return fail(nullptr, RC_ERR_INVALID, "rc_create: unknown flags");
    int ndev = 0;
    if (rc_device_count(&ndev) != RC_OK) return RC_ERR_NO_DEVICE;
    if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, RC_ERR_INVALID, "rc_create: device ordinal out of range");
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return fail(nullptr, RC_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));

    rc_ctx *c = new (std::nothrow) rc_ctx();
    if (!c) return fail(nullptr, RC_ERR_OOM, "rc_create: host allocation failed");
    c->cfg = *cfg;
    c->storage = (Storage)cfg->storage;
    const uint32_t L = cfg->levels ? cfg->levels : coverage_levels(cfg->width, cfg->height);
    c->cfg.levels = L;
    for (auto &ev : c->ev) {
        if (hipEventCreate(&ev) != hipSuccess) { rc_destroy(c); return fail(nullptr, RC_ERR_HIP, "hipEventCreate failed"); }
    }
    rc_status st = allocate(c, cfg->width, cfg->height, L);
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
    return allocate(c, W, H, L);
}

rc_status rc_run_device(rc_ctx *c, const rc_gbuffer *g, float *irr, void *stream)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_run_device: ctx is NULL");
    if (!g || !g->depth || !g->emissive || !irr) return fail(c, RC_ERR_NULL, "rc_run_device: G-buffer or output is NULL");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (graph_enabled(c)) return run_frame_graph(c, g->depth, g->emissive, irr, (hipStream_t)stream);
    return run_frame(c, g->depth, g->emissive, irr, (hipStream_t)stream);
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
    rc_status st = run_frame(c, c->g_depth, c->g_emissive, c->irr, nullptr);
    if (st != RC_OK) return st;
    HIP_TRY(c, hipMemcpy(irr_host, c->irr, 3 * (size_t)c->W * c->H * sizeof(float), hipMemcpyDeviceToHost));
    return RC_OK;
}

void rc_destroy(rc_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->cfg.device);
    hipDeviceSynchronize();
    release(c);
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
    return RC_OK;
}

rc_status rc_reset_stats(rc_ctx *c)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_reset_stats: ctx is NULL");
    for (auto &v : c->stats.stage_ms) v = 0.0;
    for (auto &v : c->stats.stage_launches) v = 0;
    c->stats.runs = 0;
    return RC_OK;
}

rc_status rc_debug_seeds(rc_ctx *c, uint32_t *dev_out, void *stream)
{
    if (!c || !dev_out) return fail(c, RC_ERR_NULL, "rc_debug_seeds: NULL argument");
    if (!c->ran || c->dbg_seed_buf < 0) return fail(c, RC_ERR_STATE, "rc_debug_seeds: needs a run with RC_FLAG_KEEP_DEBUG");
    HIP_TRY(c, hipMemcpyAsync(dev_out, c->seed[c->dbg_seed_buf], (size_t)c->W * c->H * 4, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
    return RC_OK;
}

rc_status rc_debug_sdf(rc_ctx *c, float *dev_out, void *stream)
{
    if (!c || !dev_out) return fail(c, RC_ERR_NULL, "rc_debug_sdf: NULL argument");
    if (!c->ran) return fail(c, RC_ERR_STATE, "rc_debug_sdf: no run yet");
    HIP_TRY(c, launch_sdf_untile(c->sdf, dev_out, c->W, c->H, (hipStream_t)stream));
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
    HIP_TRY(c, launch_level_to_f32(c->level[level], c->storage, c->level_texels[level], dev_out, (hipStream_t)stream));
    return RC_OK;
}

rc_status rc_debug_march(rc_ctx *c, uint32_t level, int32_t *hit_dev, int64_t *index_dev, void *stream)
{
    if (!c || !hit_dev || !index_dev) return fail(c, RC_ERR_NULL, "rc_debug_march: NULL argument");
    if ((int)level >= c->L) return fail(c, RC_ERR_INVALID, "rc_debug_march: level out of range");
    if (!c->ran) return fail(c, RC_ERR_STATE, "rc_debug_march: no run yet");
    // march-only: the probe kernel reads the distance field, never emissive
    LevelArgs a = level_args(c, (int)level, nullptr);
    HIP_TRY(c, launch_march_probe(a, c->spec, hit_dev, index_dev, (hipStream_t)stream));
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

rc_status rc_set_march(rc_ctx *c, uint32_t split, uint32_t speculation, uint32_t rays_per_lane)
{
    if (!c) return fail(nullptr, RC_ERR_NULL, "rc_set_march: ctx is NULL");
    if (split > 1 || (speculation != 1 && speculation != 2 && speculation != 4) ||
        (rays_per_lane != 1 && rays_per_lane != 2 && rays_per_lane != 4 && rays_per_lane != 8))
        return fail(c, RC_ERR_INVALID, "rc_set_march: split 0/1, speculation 1/2/4, rays_per_lane 1/2/4/8");
    c->march_split = (int)split;
    c->spec = (int)speculation;
    c->march_m = (int)rays_per_lane;
    drop_graph(c);
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
    {
        float best = 1e30f;
        int best_fuse = c->jfa_fuse;
        const int saved = c->jfa_fuse;
        for (int fuse : {0, 2, 4, 8}) {
            c->jfa_fuse = fuse;
            HIP_TRY(c, hipEventRecord(e0, s));
            for (uint32_t it = 0; it < iters; ++it) {
                rc_status r = run_frame(c, g->depth, g->emissive, irr, s, true);
                if (r != RC_OK) { c->jfa_fuse = saved; return r; }
            }
            HIP_TRY(c, hipEventRecord(e1, s));
            HIP_TRY(c, hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; best_fuse = fuse; }
        }
        c->jfa_fuse = best_fuse;
    }
    {
        float best = 1e30f;
        int best_split = c->march_split, best_spec = c->spec, best_m = c->march_m;
        for (int split : {0, 1})
        for (int spec : {1, 2, 4})
            for (int m : {1, 2, 4, 8}) {
                if (!split && (spec != 1 || m != 1)) continue;
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
        if (best_split) HIP_TRY(c, launch_march_all(march_all_args(c), c->spec, c->march_m, s));
    }
    for (int i = 1; i < c->L; ++i) {
        float best = 1e30f;
        int best_shape = c->tile[i], best_rpt = c->rpt[i];
        LevelArgs a = level_args(c, i, g->emissive);
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

}  // extern "C"
