// rc_batch.hip -- many views of uneven sizes as one work list.
//
// The reference renders one view per frame (src/window/app.rs:221-267); a
// renderer with several cameras or sub-views would run the pass once per
// view.  Here the frames of all views run as a fixed handful of launches,
// each covering every view that has work in that stage:
//
//   distance   3 launches, every view: occupancy words, band ends, row
//              lower envelopes (the exact transform, a context's default;
//              the row kernel's LDS is sized for the widest view)
//   march      1 launch of (view, level >= 1) jobs, top level first
//   levels     one launch per level index i, top down, over the views with
//              more than i levels (each view's top level has no upper level)
//   level 0    1 launch, every view (the gather writes each irradiance)
//
// Every launch is a single-view kernel body behind a job lookup (batch_job),
// so each view's irradiance equals its single-view frame bit for bit.  The
// views share one workspace allocation; the job tables are built on the host
// for the current tuning and uploaded once, and the launches are captured in
// one hipGraph, so a batch is one graph launch of ~10 kernel nodes whatever
// the number of views.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rc_batch.hpp"

namespace rc {

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

enum WlKind { WL_OCC, WL_ENDS, WL_ROWS, WL_MARCH, WL_LEVEL, WL_LEVEL0 };

struct WlView {
    int W, H, L;
    const float *depth, *emissive;
    float *irr;
    size_t bits, ends, sdf, emis4, codes[RC_MAX_LEVELS], level[RC_MAX_LEVELS];
};

struct WlLaunch {
    WlKind kind;
    int arg;               // LDS row capacity (WL_ROWS) or level (WL_LEVEL)
    size_t jobs, begins;   // byte offsets into the table buffer
    int njobs, blocks;
};

struct WorkList {
    int device = 0;
    Storage st = Storage::F32;
    std::vector<WlView> v;
    int maxL = 1;
    uint8_t *ws = nullptr;
    float2 *dirs = nullptr;
    // tuning
    int split = 1, spec = 1, m = 1;
    int tile[RC_MAX_LEVELS] = {}, rpt[RC_MAX_LEVELS] = {}, mmap[RC_MAX_LEVELS] = {};
    // launches of the current tuning
    std::vector<WlLaunch> launches;
    uint8_t *tables = nullptr;
    size_t tables_cap = 0;
    hipStream_t cap = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev[2] = {};
    size_t ws_bytes = 0;
};

#define WL_TRY(expr)                                                                \
    do {                                                                            \
        const hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                                     \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                \
            return e_;                                                              \
        }                                                                           \
    } while (0)

// Append one launch: its jobs (empty ones dropped) and their block ranges.
template <class J>
static bool add_launch(WorkList &w, std::vector<uint8_t> &blob, WlKind kind, int arg, const std::vector<J> &jobs,
                       const std::vector<int64_t> &blocks, int threads, std::string &err)
{
    std::vector<J> js;
    std::vector<int> begins;
    int64_t total = 0;
    for (size_t i = 0; i < jobs.size(); ++i) {
        if (blocks[i] <= 0) continue;
        js.push_back(jobs[i]);
        begins.push_back((int)total);
        total += blocks[i];
    }
    if (js.empty()) return true;
    if (total * threads >= (int64_t)1 << 31) {
        err = "rc_batch: a stage of this batch needs more than 2^31 threads; split the views into smaller batches";
        return false;
    }
    WlLaunch L{kind, arg, 0, 0, (int)js.size(), (int)total};
    size_t off = align256(blob.size());
    blob.resize(off + js.size() * sizeof(J));
    std::memcpy(blob.data() + off, js.data(), js.size() * sizeof(J));
    L.jobs = off;
    off = align256(blob.size());
    blob.resize(off + begins.size() * sizeof(int));
    std::memcpy(blob.data() + off, begins.data(), begins.size() * sizeof(int));
    L.begins = off;
    w.launches.push_back(L);
    return true;
}

static LevelArgs view_level_args(const WorkList &w, const WlView &x, int i)
{
    LevelArgs a{};
    a.sdf = w.ws + x.sdf;
    a.emissive = x.emissive;
    a.dirs = w.dirs + dir_table_offset(i);
    a.upper = i + 1 < x.L ? (const void *)(w.ws + x.level[i + 1]) : nullptr;
    a.out = i >= 1 ? (void *)(w.ws + x.level[i]) : nullptr;
    a.W = x.W;
    a.H = x.H;
    a.level = i;
    a.PX = grid_dim(x.W, i);
    a.PY = grid_dim(x.H, i);
    a.N = num_dirs(i);
    a.UX = grid_dim(x.W, i + 1);
    a.UY = grid_dim(x.H, i + 1);
    a.t0 = (float)interval_t0(i);
    a.t1 = (float)interval_t1(i);
    a.spacing = (float)probe_spacing(i);
    return a;
}

// As in the single-view frame: the split march merges each view's top level
// when it takes one ray per lane and the top level's lane order keeps a
// texel's 4 directions in one lane quad.  The views' top
// levels differ, so the batch asks it of every level's order.
static bool top_merged(const WorkList &w)
{
    if (!w.split || w.m != 1) return false;
    for (int i = 1; i < w.maxL; ++i)
        if (w.mmap[i] == 5) return false;
    return true;
}

// Job tables of the current tuning, uploaded to the device.
static hipError_t build_tables(WorkList &w, std::string &err)
{
    std::vector<uint8_t> blob;
    w.launches.clear();
    const int n = (int)w.v.size();
    bool ok = true;

    {  // exact distance transform: occupancy words, band ends, rows
        std::vector<OccJob> jo;
        std::vector<EndsJob> je;
        std::vector<RowJob> jr;
        std::vector<int64_t> bo, be, br;
        int wmax = 0;
        for (const WlView &x : w.v) {
            uint64_t *bits = (uint64_t *)(w.ws + x.bits);
            uint16_t *ends = (uint16_t *)(w.ws + x.ends);
            jo.push_back(OccJob{x.depth, x.emissive, bits, (float4 *)(w.ws + x.emis4), DepthConv{}, x.W, x.H,
                                occ_bx(x.W)});
            bo.push_back(occ_blocks(x.W, x.H));
            je.push_back(EndsJob{bits, ends, x.W, x.H});
            be.push_back(ends_blocks(x.W));
            jr.push_back(RowJob{bits, ends, w.ws + x.sdf, x.W, x.H});
            br.push_back(x.H);
            wmax = std::max(wmax, edt_row_wmax(x.W));
        }
        ok = ok && add_launch(w, blob, WL_OCC, 0, jo, bo, 1024, err);
        ok = ok && add_launch(w, blob, WL_ENDS, 0, je, be, kBlock, err);
        ok = ok && add_launch(w, blob, WL_ROWS, wmax, jr, br, kBlock, err);
    }

    if (w.split && ok) {  // split march: every (view, level >= 1), top level first
        std::vector<MarchJob> j;
        std::vector<int64_t> b;
        for (int i = w.maxL - 1; i >= 1; --i) {
            for (const WlView &x : w.v) {
                if (i >= x.L) continue;
                MarchLevel l{};
                l.dirs = w.dirs + dir_table_offset(i);
                l.codes = (int *)(w.ws + x.codes[i]);
                l.PX = grid_dim(x.W, i);
                l.PY = grid_dim(x.H, i);
                l.N = num_dirs(i);
                l.log2n = 2 + 2 * i;
                l.inv_px = 1.0f / (float)l.PX;
                l.t0 = (float)interval_t0(i);
                l.t1 = (float)interval_t1(i);
                l.spacing = (float)probe_spacing(i);
                l.dir_off = dir_table_offset(i);
                march_level_map(l, w.mmap[i]);
                // a view's top level is merged by the march, as in the single-view frame
                l.merged = (i == x.L - 1 && top_merged(w)) ? (void *)(w.ws + x.level[i]) : nullptr;
                l.storage = (int)w.st;
                const int64_t per = (int64_t)kMarchBlock * w.m;
                l.blocks = (int)(((int64_t)l.slots + per - 1) / per);
                j.push_back(MarchJob{w.ws + x.sdf, (const float4 *)(w.ws + x.emis4), x.W, x.H, l});
                b.push_back(l.blocks);
            }
        }
        ok = ok && add_launch(w, blob, WL_MARCH, 0, j, b, kMarchBlock, err);
    }

    for (int i = w.maxL - 1; i >= 1 && ok; --i) {
        std::vector<LevelArgs> j;
        std::vector<int64_t> b;
        for (const WlView &x : w.v) {
            if (i >= x.L) continue;
            if (i == x.L - 1 && w.split && top_merged(w)) continue;  // merged by the march
            LevelArgs a = view_level_args(w, x, i);
            a.emis4 = (const float4 *)(w.ws + x.emis4);
            if (w.split) a.codes = (const int *)(w.ws + x.codes[i]);
            j.push_back(a);
            b.push_back(level_blocks(a.PX, a.PY, a.N, w.tile[i], w.rpt[i]));
        }
        ok = ok && add_launch(w, blob, WL_LEVEL, i, j, b, kBlock, err);
    }

    if (ok) {
        std::vector<Level0Job> j;
        std::vector<int64_t> b;
        for (const WlView &x : w.v) {
            const LevelArgs a = view_level_args(w, x, 0);
            j.push_back(Level0Job{a, x.irr});
            b.push_back(level0_blocks(a.PX, a.PY));
        }
        ok = ok && add_launch(w, blob, WL_LEVEL0, 0, j, b, kBlock, err);
    }
    if (!ok) return hipErrorInvalidValue;

    WL_TRY(hipSetDevice(w.device));
    WL_TRY(hipDeviceSynchronize());  // a graph of the previous tables may still run
    if (blob.size() > w.tables_cap) {
        if (w.tables) WL_TRY(hipFree(w.tables));
        w.tables = nullptr;
        w.tables_cap = 0;
        WL_TRY(hipMalloc(&w.tables, blob.size()));
        w.tables_cap = blob.size();
    }
    WL_TRY(hipMemcpy(w.tables, blob.data(), blob.size(), hipMemcpyHostToDevice));
    if (w.gexec) WL_TRY(hipGraphExecDestroy(w.gexec));
    w.gexec = nullptr;
    return hipSuccess;
}

static hipError_t enqueue_launch(const WorkList &w, const WlLaunch &L, hipStream_t s)
{
    const BatchLaunch b{w.tables + L.jobs, (const int *)(w.tables + L.begins), L.njobs, L.blocks};
    switch (L.kind) {
    case WL_OCC: return launch_occ_batch(b, s);
    case WL_ENDS: return launch_ends_batch(b, s);
    case WL_ROWS: return launch_rows_batch(b, L.arg, s);
    case WL_MARCH: return launch_march_batch(b, w.spec, w.m, s);
    case WL_LEVEL: return launch_level_batch(b, w.st, w.tile[L.arg], w.rpt[L.arg], w.split != 0, s);
    case WL_LEVEL0: return launch_level0_batch(b, w.st, s);
    }
    return hipErrorInvalidValue;
}

static hipError_t enqueue_all(const WorkList &w, hipStream_t s)
{
    for (const WlLaunch &L : w.launches) {
        const hipError_t e = enqueue_launch(w, L, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

void worklist_destroy(WorkList *w)
{
    if (!w) return;
    hipSetDevice(w->device);
    hipDeviceSynchronize();
    if (w->gexec) hipGraphExecDestroy(w->gexec);
    if (w->cap) hipStreamDestroy(w->cap);
    for (auto e : w->ev)
        if (e) hipEventDestroy(e);
    hipFree(w->tables);
    hipFree(w->ws);
    hipFree(w->dirs);
    delete w;
}

hipError_t worklist_create(const rc_view *views, uint32_t n, const uint32_t *levels, Storage st, int device,
                           WorkList **out, std::string &err)
{
    *out = nullptr;
    WorkList *w = new (std::nothrow) WorkList();
    if (!w) {
        err = "rc_batch_create: host allocation failed";
        return hipErrorOutOfMemory;
    }
    w->device = device;
    w->st = st;
    size_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        WlView x{};
        x.W = (int)views[i].width;
        x.H = (int)views[i].height;
        x.L = (int)levels[i];
        x.depth = views[i].depth;
        x.emissive = views[i].emissive;
        x.irr = views[i].irradiance;
        const size_t px = (size_t)x.W * x.H;
        x.bits = off;
        off = align256(off + edt_bits_bytes(x.W, x.H));
        x.ends = off;
        off = align256(off + edt_ends_bytes(x.W, x.H));
        x.sdf = off;
        off = align256(off + sdf_elems(x.W, x.H));
        x.emis4 = off;
        off = align256(off + 16 * px);
        for (int l = 1; l < x.L; ++l) {
            const size_t probes = (size_t)grid_dim(x.W, l) * grid_dim(x.H, l);
            x.codes[l] = off;
            off = align256(off + probes * num_dirs(l) * sizeof(int));
            x.level[l] = off;
            off = align256(off + probes * (num_dirs(l) / 4) * texel_bytes(st));
        }
        w->maxL = std::max(w->maxL, x.L);
        w->v.push_back(x);
    }
    w->ws_bytes = off;
    for (int l = 0; l < RC_MAX_LEVELS; ++l) {
        w->tile[l] = default_level_tile(l, w->maxL);
        w->rpt[l] = 2;
    }
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc(&w->ws, off ? off : 256);
    std::vector<float2> dirs;
    for (int l = 0; l < w->maxL; ++l) {
        const int N = num_dirs(l);
        for (int k = 0; k < N; ++k) {
            const double th = 2.0 * M_PI * ((double)k + 0.5) / (double)N;
            dirs.push_back(make_float2((float)cos(th), (float)sin(th)));
        }
    }
    if (e == hipSuccess) e = hipMalloc(&w->dirs, dirs.size() * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(w->dirs, dirs.data(), dirs.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&w->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&w->ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&w->ev[1]);
    if (e != hipSuccess) {
        err = std::string("rc_batch_create: ") + hipGetErrorString(e);
        worklist_destroy(w);
        return e;
    }
    e = build_tables(*w, err);
    if (e != hipSuccess) {
        worklist_destroy(w);
        return e;
    }
    *out = w;
    return hipSuccess;
}

size_t worklist_workspace_bytes(const WorkList *w) { return w->ws_bytes; }

hipError_t worklist_run(WorkList *w, hipStream_t s, std::string &err)
{
    WL_TRY(hipSetDevice(w->device));
    if (!w->gexec) {
        WL_TRY(hipStreamBeginCapture(w->cap, hipStreamCaptureModeThreadLocal));
        const hipError_t e = enqueue_all(*w, w->cap);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(w->cap, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) hipGraphDestroy(g);
            err = std::string("rc_batch_run: capture: ") + hipGetErrorString(e != hipSuccess ? e : e2);
            return e != hipSuccess ? e : e2;
        }
        const hipError_t e3 = hipGraphInstantiate(&w->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e3 != hipSuccess) {
            w->gexec = nullptr;
            err = std::string("rc_batch_run: hipGraphInstantiate: ") + hipGetErrorString(e3);
            return e3;
        }
    }
    WL_TRY(hipGraphLaunch(w->gexec, s));
    return hipSuccess;
}

// Mean time of the launches `pick` selects, each run `iters` times in order
// (their inputs are those of the last whole frame).
template <class Pick>
static hipError_t time_launches(WorkList &w, hipStream_t s, int iters, Pick pick, float &ms, std::string &err)
{
    WL_TRY(hipEventRecord(w.ev[0], s));
    for (int it = 0; it < iters; ++it)
        for (const WlLaunch &L : w.launches)
            if (pick(L)) WL_TRY(enqueue_launch(w, L, s));
    WL_TRY(hipEventRecord(w.ev[1], s));
    WL_TRY(hipEventSynchronize(w.ev[1]));
    WL_TRY(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
    ms /= (float)iters;
    return hipSuccess;
}

// Stage by stage: the march's speculation depth, rays per lane and per-level
// lane order, then per level the tile shape and probes per thread,
// for the split march and for marching in the level kernels; the split form
// is kept if its march plus levels is faster.
hipError_t worklist_autotune(WorkList *w, int iters, hipStream_t s, std::string &err)
{
    WorkList &x = *w;
    if (iters < 1) iters = 1;
    auto frame = [&]() -> hipError_t {
        WL_TRY(build_tables(x, err));
        WL_TRY(enqueue_all(x, s));
        WL_TRY(hipStreamSynchronize(s));
        return hipSuccess;
    };
    float best = 1e30f, ms = 0.f;

    auto tune_levels = [&](float &total) -> hipError_t {
        total = 0.f;
        for (int i = x.maxL - 1; i >= 1; --i) {
            float lbest = 1e30f;
            int bshape = x.tile[i], brpt = x.rpt[i];
            for (int shape = 0; shape < TILE_COUNT; ++shape) {
                if (tile_kd(shape) > num_dirs(i)) continue;
                for (int r : {1, 2, 4}) {
                    x.tile[i] = shape;
                    x.rpt[i] = r;
                    WL_TRY(build_tables(x, err));
                    WL_TRY(time_launches(x, s, iters, [i](const WlLaunch &L) { return L.kind == WL_LEVEL && L.arg == i; },
                                         ms, err));
                    if (ms < lbest) lbest = ms, bshape = shape, brpt = r;
                }
            }
            x.tile[i] = bshape;
            x.rpt[i] = brpt;
            total += lbest;
        }
        return hipSuccess;
    };

    float split_total = 1e30f, inplace_total = 1e30f;
    int tiles_split[RC_MAX_LEVELS], rpts_split[RC_MAX_LEVELS];
    if (x.maxL > 1) {
        x.split = 1;
        best = 1e30f;
        int bspec = 1, bm = 1;
        for (auto sm : {std::pair<int, int>{1, 1}, {1, 2}, {2, 1}, {2, 2}, {1, 4}}) {
            x.spec = sm.first;
            x.m = sm.second;
            WL_TRY(frame());
            WL_TRY(time_launches(x, s, iters, [](const WlLaunch &L) { return L.kind == WL_MARCH; }, ms, err));
            if (ms < best) best = ms, bspec = sm.first, bm = sm.second;
        }
        x.spec = bspec;
        x.m = bm;
        for (int i = x.maxL - 1; i >= 1; --i) {  // each level's lane order, the march timed whole
            best = 1e30f;
            int bmap = x.mmap[i];
            for (int map = 0; map < kMarchMaps; ++map) {
                if (!march_map_legal(i, map)) continue;
                x.mmap[i] = map;
                WL_TRY(build_tables(x, err));
                WL_TRY(time_launches(x, s, iters, [](const WlLaunch &L) { return L.kind == WL_MARCH; }, ms, err));
                if (ms < best) best = ms, bmap = map;
            }
            x.mmap[i] = bmap;
        }
        WL_TRY(frame());
        float lv = 0.f;
        WL_TRY(tune_levels(lv));
        split_total = best + lv;
        std::memcpy(tiles_split, x.tile, sizeof(tiles_split));
        std::memcpy(rpts_split, x.rpt, sizeof(rpts_split));
        x.split = 0;
        WL_TRY(frame());
        WL_TRY(tune_levels(inplace_total));
        if (split_total <= inplace_total) {
            x.split = 1;
            std::memcpy(x.tile, tiles_split, sizeof(tiles_split));
            std::memcpy(x.rpt, rpts_split, sizeof(rpts_split));
        }
    }
    WL_TRY(build_tables(x, err));
    return hipSuccess;
}

int worklist_launches(const WorkList *w) { return (int)w->launches.size(); }

}  // namespace rc
