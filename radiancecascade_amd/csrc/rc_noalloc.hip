// rc_noalloc.hip -- rc_run_device_noalloc: the plain GI frame of
// rc_run_device without a context and without any device allocation, for
// callers that cannot allocate (code under hipGraph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// The frame's working buffers -- the exact transform's occupancy words and
// band ends, the distance field and the merged levels 1..L-1 --
// live in one file-scope device array,
// `scratch`, of kNoAllocScratchBytes; the per-level direction tables in a
// second one, `scratch_dirs`.  Each GPU the library runs on holds its own copy
// of both until the process ends.  In host code the arrays' names are
// host-side stand-ins, not device addresses: the first call on a device (with
// that device current) resolves their device addresses with
// hipGetSymbolAddress, once per device, and every launch uses those.  The
// kernels are the library's own (the exact transform's occupancy, column and
// row kernels -- the default distance field of a context -- the level kernels
// marching in place, level 0 + gather), so the irradiance equals
// rc_run_device's bit for bit.
//
// Every call on a GPU shares that GPU's scratch: calls that overlap on one GPU
// (two streams, two threads) race and give wrong results.  Nothing carried
// between calls needs clearing: every buffer the frame reads is written
// earlier in the same frame.
#include <mutex>
#include <string>
#include <vector>

#include <math.h>

#include "rc_internal.hpp"

namespace rc {

// 64 MiB: enough for a 1920x1080 frame with 6 levels (48.1 MB).
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
constexpr int kNoAllocMaxLevels = 7;
constexpr int kNoAllocDirs = 4 * ((1 << (2 * kNoAllocMaxLevels)) - 1) / 3;  // sum of N_i, i < 7

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
static __device__ __align__(256) float2 scratch_dirs[kNoAllocDirs];

constexpr int kNoAllocMaxDevices = 64;
struct NoAllocDevice {
    bool ready = false;
    uint8_t *scratch = nullptr;  // device address of `scratch`
    float2 *dirs = nullptr;      // device address of `scratch_dirs`
};
static NoAllocDevice g_noalloc[kNoAllocMaxDevices];
static std::mutex g_noalloc_mu;

struct NoAllocLayout {
    size_t bits, ends, sdf, level[RC_MAX_LEVELS], total;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static NoAllocLayout noalloc_layout(int W, int H, int L)
{
    NoAllocLayout m{};
    size_t off = 0;
    m.bits = off;
    off = align256(off + edt_bits_bytes(W, H));
    m.ends = off;
    off = align256(off + edt_ends_bytes(W, H));
    m.sdf = off;
    off = align256(off + sdf_elems(W, H));
    for (int i = 1; i < L; ++i) {
        m.level[i] = off;
        off = align256(off + (size_t)grid_dim(W, i) * grid_dim(H, i) * (num_dirs(i) / 4) * sizeof(float4));
    }
    m.total = off;
    return m;
}


// Once per device, with that device current: the device addresses of the
// two arrays, and the direction tables (theta = 2 pi (k + 1/2) / N in double,
// rounded to fp32, as the context path builds them).
static hipError_t noalloc_setup(int dev, NoAllocDevice **out)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDevice &d = g_noalloc[dev];
    if (!d.ready) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        d.scratch = (uint8_t *)p;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch_dirs));
        if (e != hipSuccess) return e;
        d.dirs = (float2 *)p;
        std::vector<float2> dirs;
        dirs.reserve(kNoAllocDirs);
        for (int i = 0; i < kNoAllocMaxLevels; ++i) {
            const int N = num_dirs(i);
            for (int k = 0; k < N; ++k) {
                const double th = 2.0 * M_PI * ((double)k + 0.5) / (double)N;
                dirs.push_back(make_float2((float)cos(th), (float)sin(th)));
            }
        }
        e = hipMemcpy(d.dirs, dirs.data(), dirs.size() * sizeof(float2), hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
        d.ready = true;
    }
    *out = &d;
    return hipSuccess;
}

}  // namespace rc

using namespace rc;

static thread_local std::string g_noalloc_err;

static rc_status nfail(rc_status s, const std::string &msg)
{
    g_noalloc_err = msg;
    return s;
}

extern "C" {

const char *rc_noalloc_last_error(void) { return g_noalloc_err.c_str(); }

uint64_t rc_noalloc_scratch_bytes(uint32_t width, uint32_t height, uint32_t levels)
{
    if (width < 1 || height < 1 || width > RC_MAX_DIM || height > RC_MAX_DIM || levels < 1 ||
        levels > (uint32_t)kNoAllocMaxLevels)
        return 0;
    return noalloc_layout((int)width, (int)height, (int)levels).total;
}

rc_status rc_run_device_noalloc(uint32_t width, uint32_t height, uint32_t levels, const rc_gbuffer *g, float *irr,
                                void *stream)
{
    if (!g || !g->depth || !g->emissive || !irr) return nfail(RC_ERR_NULL, "rc_run_device_noalloc: NULL argument");
    if (width < 1 || height < 1 || width > RC_MAX_DIM || height > RC_MAX_DIM)
        return nfail(RC_ERR_INVALID, "rc_run_device_noalloc: width/height out of range");
    if (levels < 1 || levels > (uint32_t)kNoAllocMaxLevels)
        return nfail(RC_ERR_INVALID, "rc_run_device_noalloc: levels must be 1.." + std::to_string(kNoAllocMaxLevels));
    const int W = (int)width, H = (int)height, L = (int)levels;
    const NoAllocLayout lay = noalloc_layout(W, H, L);
    if (lay.total > kNoAllocScratchBytes)
        return nfail(RC_ERR_INVALID, "rc_run_device_noalloc: " + std::to_string(W) + "x" + std::to_string(H) + " with " +
                                         std::to_string(L) + " levels needs " + std::to_string(lay.total) +
                                         " bytes of scratch; the static scratch holds " +
                                         std::to_string(kNoAllocScratchBytes));
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return nfail(RC_ERR_NO_DEVICE, std::string("rc_run_device_noalloc: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return nfail(RC_ERR_INVALID, "rc_run_device_noalloc: device ordinal too large");
    NoAllocDevice *d = nullptr;
    e = noalloc_setup(dev, &d);
    if (e != hipSuccess)
        return nfail(RC_ERR_HIP, std::string("rc_run_device_noalloc: set-up: ") + hipGetErrorString(e));

    const hipStream_t s = (hipStream_t)stream;
    uint8_t *sdf = d->scratch + lay.sdf;
    const DepthConv linear{};
#define NTRY(expr)                                                                                              \
    do {                                                                                                        \
        const hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return nfail(RC_ERR_HIP, std::string("rc_run_device_noalloc: ") + #expr + ": " + \
                                                           hipGetErrorString(e_));                             \
    } while (0)
    NTRY(launch_edt(g->depth, linear, g->emissive, (uint64_t *)(d->scratch + lay.bits),
                    (uint16_t *)(d->scratch + lay.ends), sdf, nullptr, W, H, s));
    for (int i = L - 1; i >= 0; --i) {
        LevelArgs a{};
        a.sdf = sdf;
        a.emissive = g->emissive;
        a.emis4 = nullptr;
        a.dirs = d->dirs + dir_table_offset(i);
        a.upper = i + 1 < L ? (const void *)(d->scratch + lay.level[i + 1]) : nullptr;
        a.out = i >= 1 ? (void *)(d->scratch + lay.level[i]) : nullptr;
        a.W = W;
        a.H = H;
        a.level = i;
        a.PX = grid_dim(W, i);
        a.PY = grid_dim(H, i);
        a.N = num_dirs(i);
        a.UX = grid_dim(W, i + 1);
        a.UY = grid_dim(H, i + 1);
        a.t0 = (float)interval_t0(i);
        a.t1 = (float)interval_t1(i);
        a.spacing = (float)probe_spacing(i);
        if (i >= 1) {
            // the context path's default tiling (the result does not depend on it)
            NTRY(launch_level(a, Storage::F32, default_level_tile(i, L), 2, s));
        } else {
            NTRY(launch_level0_gather(a, Storage::F32, irr, nullptr, s));
        }
    }
#undef NTRY
    return RC_OK;
}

}  // extern "C"
