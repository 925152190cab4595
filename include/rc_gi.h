/*
 * rc_gi.h -- C ABI of the MI355X radiance-cascade GI pass (librcgi.so).
 *
 * The reference renderer (jw910731/RadianceCascade) exposes one pass
 * interface, the Rust trait
 *
 *     pub trait RenderStage<T> {                                src/app.rs:3-7
 *         fn render(&self, state: &mut T, view: &TextureView, encoder: &mut CommandEncoder);
 *         fn resize(&mut self, device: &Device, config: &SurfaceConfiguration);
 *         fn update(&mut self, state: &T, queue: &Queue);
 *     }
 *
 * implemented by DefaultRenderer (src/renderer.rs:558-632) and driven once per
 * frame by AppInternal::handle_redraw (src/window/app.rs:221-267) in the order
 * update -> render -> submit.  The entry points below are what a Rust FFI
 * crate implementing RenderStage<AppState> for a GI stage binds:
 *
 *   rc_create      <- DefaultRenderer::new             src/renderer.rs:168-555
 *   rc_resize      <- RenderStage::resize              src/app.rs:5, src/renderer.rs:615-618
 *   rc_update_host <- RenderStage::update (uploads)    src/app.rs:6, src/renderer.rs:620-631
 *   rc_render_host <- RenderStage::render              src/app.rs:4, src/renderer.rs:559-613
 *   rc_run_device  <- RenderStage::render, device-resident G-buffer, stream ordered
 *   rc_destroy     <- Drop of the stage (wgpu resources, src/renderer.rs:156-165)
 *
 * The reference has no GI pass at all (its only pass is the Blinn-Phong
 * forward raster); the maths is the published radiance-cascade method.
 *
 * Conventions: plain-old-data structs carry a version field; every call
 * returns an rc_status and never throws; the context owns all device memory it
 * allocates, the caller owns the buffers it passes; one context is used by one
 * thread at a time and its device calls are ordered on the stream given (calls
 * on one context must not overlap on different streams).  Device pointers are
 * HIP device pointers; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).
 */
#ifndef RC_GI_H
#define RC_GI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RC_ABI_VERSION 2u  /* 2: rc_config carries the depth convention */
#define RC_MAX_LEVELS 8
#define RC_MAX_DIM 16384
#define RC_MAX_BOXES 256

typedef enum {
    RC_OK = 0,
    RC_ERR_INVALID = 1,     /* bad dimensions, level count, version, storage */
    RC_ERR_NULL = 2,        /* a required pointer was NULL */
    RC_ERR_NO_DEVICE = 3,   /* no HIP device visible */
    RC_ERR_HIP = 4,         /* a HIP runtime call failed (see rc_last_error) */
    RC_ERR_OOM = 5,         /* device allocation failed */
    RC_ERR_STATE = 6        /* call out of order (e.g. render before update) */
} rc_status;

/* Storage precision of the materialised merged cascade levels.  Marching,
 * merging and the gather always accumulate in fp32. */
typedef enum { RC_STORAGE_F32 = 0, RC_STORAGE_F16 = 1, RC_STORAGE_BF16 = 2 } rc_storage;

/* rc_config.flags */
#define RC_FLAG_PROFILE    0x1u  /* time every stage with HIP events (rc_get_stats) */
#define RC_FLAG_KEEP_DEBUG 0x2u  /* also write the jump-flood seed field (that mode) and level 0 for rc_debug_* */
#define RC_FLAG_GRAPH      0x8u  /* rc_run_device replays the frame as one captured hipGraph
                                    (one host call per frame) instead of launching its
                                    kernels one by one; on MI355X each replay leaves ~4 us
                                    between frames, so direct launches are faster whenever
                                    the host runs ahead of the GPU */

/* rc_config.depth_mode: what the depth plane holds.
 *   RC_DEPTH_LINEAR: linear view depth D (distance along the optical axis).
 *   RC_DEPTH_NDC:    the reference renderer's Depth32Float value, NDC depth d
 *                    in [0,1] of glam's perspective_rh (src/camera.rs:77-79,
 *                    src/texture.rs:149): d = f (D - n) / ((f - n) D) for the
 *                    near / far planes n = z_near, f = z_far (the reference
 *                    uses 0.1 and 100: Projection::new at src/app.rs:26).
 *                    The pass converts every depth it reads to
 *                    D = (n f) / (f - d (f - n)) in fp32, n f and f - n
 *                    rounded to fp32 first, then applies the linear rule. */
#define RC_DEPTH_LINEAR 0u
#define RC_DEPTH_NDC    1u

typedef struct {
    uint32_t version;   /* RC_ABI_VERSION */
    uint32_t width;     /* 1 .. RC_MAX_DIM */
    uint32_t height;    /* 1 .. RC_MAX_DIM */
    uint32_t levels;    /* 1 .. RC_MAX_LEVELS; 0 = enough to cover the diagonal */
    uint32_t storage;   /* rc_storage */
    uint32_t flags;     /* RC_FLAG_* */
    int32_t device;     /* HIP device ordinal */
    uint32_t depth_mode;  /* RC_DEPTH_LINEAR or RC_DEPTH_NDC */
    float z_near;       /* RC_DEPTH_NDC: near plane, 0 < z_near < z_far (ignored for linear depth) */
    float z_far;        /* RC_DEPTH_NDC: far plane, finite */
} rc_config;

/* Planar fp32 G-buffer: depth [H][W] (rc_config.depth_mode), emissive [3][H][W]
 * (radiance), albedo [3][H][W].  Albedo is not read by the GI pass (GI is
 * emission transport; compositing multiplies by albedo outside the pass) and
 * may be NULL. */
typedef struct {
    const float *depth;
    const float *albedo;
    const float *emissive;
} rc_gbuffer;

/* Per-stage device times in milliseconds, summed over the runs since the last
 * rc_reset_stats (only with RC_FLAG_PROFILE), and the run count. */
#define RC_STAGE_SEED 0         /* occupancy: exact mode's column words, or the jump flood's seeds */
#define RC_STAGE_JFA 1          /* the distance transform: exact mode's column + row kernels, or all global jump-flood passes */
#define RC_STAGE_JFA_TAIL 2     /* fused short-step passes + distance field */
#define RC_STAGE_LEVEL0 3       /* RC_STAGE_LEVEL0 + i = level i march/merge; level 0 includes the gather */
#define RC_STAGE_MARCH (RC_STAGE_LEVEL0 + RC_MAX_LEVELS) /* split march of all levels >= 1 */
/* learned far field (rc_set_surrogate mode 1) */
#define RC_STAGE_NEAR_SDF (RC_STAGE_MARCH + 1)  /* exact distance field up to 10 px */
#define RC_STAGE_MIP (RC_STAGE_MARCH + 2)       /* emitter / occluder mip pyramid */
#define RC_STAGE_CONE (RC_STAGE_MARCH + 3)      /* cone trace (level-2 and level-3 probe grids) */
#define RC_STAGE_SURROGATE (RC_STAGE_MARCH + 4) /* the network's level 2 (features + MLP) */
#define RC_STAGE_FALLBACK (RC_STAGE_MARCH + 5)  /* plain far field + patch of flagged tiles */
#define RC_NUM_STAGES (RC_STAGE_MARCH + 6)
typedef struct {
    uint32_t version;
    uint32_t runs;
    double stage_ms[RC_NUM_STAGES];
    uint32_t stage_launches[RC_NUM_STAGES];
    uint64_t level_rays[RC_MAX_LEVELS];
    uint64_t level_bytes_stored[RC_MAX_LEVELS]; /* bytes of merged level i written per run */
    /* learned far field: frames run, frames that fell back, flagged tiles
     * (summed over frames) and tiles per frame */
    uint64_t surrogate_frames;
    uint64_t fallback_frames;
    uint64_t flagged_tiles;
    uint64_t tiles_per_frame;
} rc_stats;

typedef struct rc_ctx rc_ctx;

/* ---- the pass (RenderStage mirror) -------------------------------------- */
rc_status rc_create(const rc_config *cfg, rc_ctx **out);
rc_status rc_resize(rc_ctx *ctx, uint32_t width, uint32_t height, uint32_t levels);
rc_status rc_update_host(rc_ctx *ctx, const rc_gbuffer *host);
rc_status rc_render_host(rc_ctx *ctx, float *irradiance_host);
/* Page-locked host memory for the host-staging path: G-buffer planes and
 * irradiance held in it move by DMA at the link's rate, where pageable memory
 * goes through the runtime's staging copies.  No counterpart in the
 * reference; its uploads (src/renderer.rs:620-631) would read their source
 * planes from such buffers.  rc_host_free(NULL) is a no-op. */
rc_status rc_host_alloc(uint64_t bytes, void **out);
void rc_host_free(void *ptr);
rc_status rc_run_device(rc_ctx *ctx, const rc_gbuffer *dev, float *irradiance_dev, void *stream);

/* ---- allocation-free frame -----------------------------------------------------
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The plain GI frame of rc_run_device --
 * linear depth, fp32 storage; the irradiance is equal bit for bit -- with no
 * context: its working buffers (jump-flood fields, distance field, merged
 * levels) live in a static 64 MiB device array that each GPU holds until the
 * process ends (enough for 1920x1080 with 6 levels; rc_noalloc_scratch_bytes
 * gives a size's need, larger sizes are refused with RC_ERR_INVALID), and
 * its direction tables in a static array set on the first call on a device.
 * It calls no hipMalloc, hipFree or other allocating routine.  Runs on the
 * current device, stream ordered.  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (several streams or threads) race.
 * levels: 1..7.  Errors: rc_noalloc_last_error. */
rc_status rc_run_device_noalloc(uint32_t width, uint32_t height, uint32_t levels, const rc_gbuffer *dev,
                                float *irradiance_dev, void *stream);
uint64_t rc_noalloc_scratch_bytes(uint32_t width, uint32_t height, uint32_t levels);  /* 0 if out of range */
const char *rc_noalloc_last_error(void);
void rc_destroy(rc_ctx *ctx);

/* ---- level-kernel tiling (autotuner) ----------------------------------------
 * A level-i kernel (i >= 1) runs workgroups of TX x TY probes x KD directions
 * (TX*TY*KD = 256 lanes) with R rays per lane (probe rows TY apart).
 * shape: 0 = 8x8x4, 1 = 4x4x16, 2 = 2x2x64, 3 = 1x1x256, 4 = 8x4x8 (TX x TY x KD;
 * KD <= N_i required);
 * rays_per_lane: 1, 2 or 4.  Every tiling computes identical results.
 * rc_autotune times every legal (shape, rays_per_lane) of every level on the
 * given device G-buffer (iters launches each, HIP events) and keeps the
 * fastest; rc_get_level_tiling reports the current choice. */
#define RC_NUM_TILE_SHAPES 5
rc_status rc_set_level_tiling(rc_ctx *ctx, uint32_t level, uint32_t shape, uint32_t rays_per_lane);
rc_status rc_get_level_tiling(const rc_ctx *ctx, uint32_t level, uint32_t *shape, uint32_t *rays_per_lane);
rc_status rc_autotune(rc_ctx *ctx, const rc_gbuffer *dev, float *irradiance_dev, uint32_t iters, void *stream);
/* Distance transform of the SDF (no counterpart in the reference, which has
 * no distance field).  RC_SDF_EXACT (the default): the exact squared
 * Euclidean distance to the nearest occupied pixel, by an occupancy kernel,
 * a column kernel and a row kernel (3 launches).  RC_SDF_JUMP_FLOOD: the jump
 * flood of ceil(log2 max(W, H)) passes, whose distance can exceed the exact
 * one at a few pixels (< 1 %; never below it); its seed field is readable
 * with rc_debug_seeds.  Both fields are stored as the same one-byte code.
 * Changing the mode allocates the mode's buffers if needed. */
#define RC_SDF_EXACT 0u
#define RC_SDF_JUMP_FLOOD 1u
rc_status rc_set_sdf_mode(rc_ctx *ctx, uint32_t mode);
uint32_t rc_get_sdf_mode(const rc_ctx *ctx);
/* Jump-flood mode: passes with step <= fused_step (0, 1, 2, 4 or 8) run fused
 * in one LDS-tiled kernel; 0 runs every pass as a global kernel.  The result
 * is identical; rc_autotune also picks this. */
rc_status rc_set_jfa_fuse(rc_ctx *ctx, uint32_t fused_step);
uint32_t rc_get_jfa_fuse(const rc_ctx *ctx);
/* Image rows per wave of a global jump-flood pass: 1, 2 or 4 (no counterpart
 * in the reference).  More rows share a wave's row setup and keep more loads
 * in flight per lane; the result is identical and rc_autotune picks it. */
rc_status rc_set_jfa_rows(rc_ctx *ctx, uint32_t rows);
uint32_t rc_get_jfa_rows(const rc_ctx *ctx);
/* split = 1: every level >= 1 is marched by one launch (top level first, so
 * the longest rays start earliest and the levels' dependent chains overlap),
 * then the levels are merged top-down from the stored ray codes; split = 2:
 * the same march by a persistent grid of resident waves, each walking an
 * interleaved share of 64-ray packets and handing a new ray to a lane as soon
 * as its ray ends (rays_per_lane must be 1); split = 0: each level kernel
 * marches and merges its own rays.  speculation (1, 2, 4) is the split
 * march's look-ahead depth; rays_per_lane (1, 2, 4, 8) how many rays a lane
 * of the split = 1 march takes one after another.  The result is identical;
 * rc_autotune also picks these (a one-level frame has nothing to march). */
rc_status rc_set_march(rc_ctx *ctx, uint32_t split, uint32_t speculation, uint32_t rays_per_lane);
rc_status rc_get_march(const rc_ctx *ctx, uint32_t *split, uint32_t *speculation, uint32_t *rays_per_lane);
/* Order of a level's rays over the lanes of the split march (split = 1; no
 * counterpart in the reference; every order stores the same codes): 0 =
 * canonical order (a wave marches 64 consecutive rays: 64 directions of one
 * probe, or 64/N whole probes); wave tiles of probes x directions, 1 = 2x2 x 16,
 * 2 = 4x4 x 4, 3 = 4x2 x 8, 4 = 2x1 x 32, 5 = 8x8 x 1 (the slice at most
 * N_level).  Rays from nearby probes in nearby directions march for similar
 * lengths, so fewer lanes idle; rc_autotune picks the order per level. */
rc_status rc_set_march_order(rc_ctx *ctx, uint32_t level, uint32_t order);
rc_status rc_get_march_order(const rc_ctx *ctx, uint32_t level, uint32_t *order);

/* ---- introspection -------------------------------------------------------- */
rc_status rc_get_config(const rc_ctx *ctx, rc_config *out);
rc_status rc_get_stats(const rc_ctx *ctx, rc_stats *out);
rc_status rc_reset_stats(rc_ctx *ctx);
const char *rc_last_error(const rc_ctx *ctx);   /* ctx may be NULL: last global error */
const char *rc_status_string(rc_status s);
uint32_t rc_abi_version(void);
rc_status rc_device_count(int *count);

/* Parity readbacks of the last run (device destination, stream ordered).
 * rc_debug_seeds / rc_debug_level(.., 0, ..) need RC_FLAG_KEEP_DEBUG (and
 * rc_debug_seeds the jump-flood mode).  After a
 * learned frame (rc_set_surrogate mode 1) rc_debug_level returns that frame's
 * levels: 2 the network's, 1 merged from it (with flagged tiles patched), 0
 * the fluence; it has no level above 2 (RC_ERR_STATE).
 *   seeds : uint32 [H][W]  final jump-flood field, (y<<16)|x or 0xFFFFFFFF
 *   sdf   : float  [H][W]  distance values the march reads, in px (2160 with
 *           no occupied pixel)
 *   level : float  [Py_i][Px_i][max(N_i/4,1)][4], merged level i averaged over
 *           the 4 children of each level-(i-1) direction (fp32 whatever the
 *           storage; level 0 is the per-probe fluence) */
rc_status rc_debug_seeds(rc_ctx *ctx, uint32_t *dev_out, void *stream);
rc_status rc_debug_sdf(rc_ctx *ctx, float *dev_out, void *stream);
rc_status rc_debug_level(rc_ctx *ctx, uint32_t level, float *dev_out, void *stream);
/* The ray codes the last frame's kernels stored for one level, per ray in
 * canonical order ((py*Px+px)*N+k): the hit pixel y*W+x, -1 (left the image)
 * or -2 (reached the interval end).  Levels marched by the split march
 * (rc_set_march split=1) return its code buffer, read by the merge kernels;
 * levels marched in place (split=0, level 0, and levels 0-1 of the learned
 * frame) return the copy those kernels write under RC_FLAG_KEEP_DEBUG
 * (RC_ERR_STATE otherwise, and for levels the learned frame does not march).
 * index_dev (optional, levels >= 1): per canonical ray, the slot of the
 * split march (with the level's current lane order) whose decode into (probe
 * column, row, direction) lands on that ray, -1 if none: a permutation of
 * the slots when every ray is marched exactly once (the identity in the
 * canonical order). */
rc_status rc_debug_march(rc_ctx *ctx, uint32_t level, int32_t *hit_dev, int64_t *index_dev, void *stream);

/* ---- error norms (FP64 reductions on the GPU) ------------------------------
 * x: n fp32 values; ref: n values, fp64 (ref_is_f64 = 1) or fp32.  out_dev
 * (3 doubles, device): sum (x-ref)^2, sum ref^2, max |x-ref|; rel-L2 =
 * sqrt(out[0]/out[1]).  Deterministic (fixed two-stage order).  Uses the
 * context's workspace, so calls on one context are stream ordered. */
rc_status rc_error_norms(rc_ctx *ctx, const float *x, const void *ref, uint32_t ref_is_f64, uint64_t n,
                         double *out_dev, void *stream);

/* ---- composite (the presented frame) -------------------------------------------
 * Replaces what the reference renderer presents after its one pass
 * (DefaultRenderer::render, src/renderer.rs:559-613, into the Bgra8UnormSrgb
 * surface of src/window/app.rs:59), with GI added: per pixel and channel,
 * linear c = (albedo * irradiance + emissive) + direct in fp32 (direct: the
 * forward pass's Blinn-Phong term of src/shader.wgsl:76-101, planar fp32
 * [3][H][W], or NULL for none), encoded to 8-bit sRGB (round to nearest of
 * the IEC 61966-2-1 code via the thresholds of rc_srgb_thresholds; c <= 0 ->
 * 0, c >= 1 -> 255) and written as one uint32 per pixel holding the bytes
 * B, G, R, A = 255.  albedo, emissive, irradiance: planar fp32 [3][H][W] at
 * the context's size; all device pointers. */
rc_status rc_composite(rc_ctx *ctx, const float *albedo, const float *emissive, const float *irradiance,
                       const float *direct, uint32_t *bgra8, void *stream);
/* The 256 encoding thresholds (host array): t[k], k = 1..255, is the linear
 * value whose sRGB code is (k - 1/2) / 255 (double, rounded to fp32); t[0] is
 * -inf.  The 8-bit value of c is the largest k with c >= t[k]. */
void rc_srgb_thresholds(float *t);

/* ---- dump files (host only) ----------------------------------------------------
 * No counterpart in the reference (it keeps no GI data on disk).  An
 * "RCDUMP01" file holds named planar fp32 arrays [channels][height][width]
 * (G-buffer planes, merged cascades, irradiance) with a SHA-256 of its plane
 * records and payload: fixtures that travel between machines and captured
 * frames for the trainer.  Readers check magic, version, that the sizes match
 * the file, and the checksum (any mismatch: RC_ERR_INVALID, message in
 * rc_dump_last_error). */
typedef struct {
    char name[16];                   /* NUL-padded; unique within a file */
    uint32_t channels, height, width, reserved;
    const float *data;               /* host values (write); NULL from rc_dump_info */
} rc_plane;
rc_status rc_dump_write(const char *path, const rc_plane *planes, uint32_t n_planes);
/* names and shapes of up to `cap` planes; *n_out = the file's plane count */
rc_status rc_dump_info(const char *path, rc_plane *planes, uint32_t cap, uint32_t *n_out);
/* copies plane `name` (count = channels*height*width floats) into dst */
rc_status rc_dump_read(const char *path, const char *name, float *dst, uint64_t count);
const char *rc_dump_last_error(void);
/* SHA-256 (FIPS 180-4) of n bytes into out32[32] */
void rc_sha256(const void *data, uint64_t n, uint8_t *out32);

/* ---- learned far field ------------------------------------------------------
 * No counterpart in the reference (it has no GI).  Mode 1 replaces levels >= 2
 * of the frame: a cone trace through an emitter / occluder mip pyramid gives,
 * per level-2 probe and level-1 direction k, a far-field estimate c_k; an MLP
 * (64 -> 64 -> 64 -> 64 -> 64, ReLU, BF16 weights on MFMA, FP32 accumulation)
 * maps x = standardised (log(1 + c_k) rgb, cone transmittance T_k) for the 16
 * directions to z, and the merged level 2 is
 * A_2 = max(c_k + out_scale * z_k, 0); levels 1 and 0 are then the plain
 * kernels (marched on a distance field exact up to 10 px).  Output 48
 * predicts log of the squared error.  Tiles of 8x8 level-2 probes whose mean
 * predicted error exceeds `threshold` take the plain cascades' level 1
 * (fallback: one full plain frame plus a patch, decided on the host after the
 * frame).  Needs levels >= 3. */
#define RC_MLP_INPUTS 64     /* per direction k: log(1 + c_k) rgb, cone transmittance T_k */
#define RC_MLP_OUTPUTS 64
#define RC_MLP_PARAMS 16640  /* for l = 0..3: W_l [K_l][64] (K = 64) then b_l [64] */
#define RC_SURROGATE_BLOB_MAGIC "RCSURR01"
typedef struct {
    uint32_t version;   /* RC_ABI_VERSION */
    uint32_t mode;      /* 0 plain cascades, 1 learned far field */
    float threshold;    /* fallback threshold (normalised squared error); +inf: never */
    uint32_t reserved;
} rc_surrogate_config;
rc_status rc_set_surrogate(rc_ctx *ctx, const rc_surrogate_config *cfg);
rc_status rc_get_surrogate(const rc_ctx *ctx, rc_surrogate_config *out);
/* Network parameters (host arrays): params RC_MLP_PARAMS fp32; in_mean,
 * in_istd RC_MLP_INPUTS each (x = (feature - mean) * istd). */
rc_status rc_mlp_set(rc_ctx *ctx, const float *params, const float *in_mean, const float *in_istd, float out_scale);
rc_status rc_mlp_get(const rc_ctx *ctx, float *params, float *in_mean, float *in_istd, float *out_scale);
/* Serialised weights: "RCSURR01", uint32 version, inputs, outputs, params,
 * float out_scale, uint32 train_scenes, then in_mean, in_istd, params (fp32
 * little endian).  size = bytes needed (call with buf NULL to query). */
rc_status rc_surrogate_save(const rc_ctx *ctx, void *buf, uint64_t capacity, uint64_t *size);
rc_status rc_surrogate_load(rc_ctx *ctx, const void *buf, uint64_t size);
/* The network's kernels on device buffers (parity hooks for the training and
 * inference kernels).  feats: [n][RC_MLP_INPUTS] features (standardised inside);
 * out: [n][64] raw outputs z.  rc_mlp_gradient: targets [n][48] normalised
 * residuals; grad: RC_MLP_PARAMS fp32 of the batch-mean loss
 * L = sum (z - t)^2 / (48 n) + 0.05 sum (z_48 - log(mse_s + 1e-2))^2 / n;
 * loss: 1 float.  rc_mlp_adam: one Adam step (0.9, 0.999, 1e-8) with grad. */
rc_status rc_mlp_forward(rc_ctx *ctx, const float *feats_dev, uint32_t n, float *out_dev, void *stream);
rc_status rc_mlp_gradient(rc_ctx *ctx, const float *feats_dev, const float *targets_dev, uint32_t n, float *grad_dev,
                          float *loss_dev, void *stream);
rc_status rc_mlp_adam(rc_ctx *ctx, const float *grad_dev, float lr, void *stream);
/* Training rows of one G-buffer at the context's size: runs the plain frame,
 * then per level-2 probe writes feats [P][RC_MLP_INPUTS] = per direction k
 * (log(1 + c_k) rgb, T_k) and targets [P][48] = A_2(plain) - c
 * (unnormalised), P = Px_2 * Py_2. */
rc_status rc_surrogate_samples(rc_ctx *ctx, const rc_gbuffer *dev, float *feats_dev, float *targets_dev,
                               void *stream);
/* Train the context's network on the GPU on plain-path outputs of seeded
 * scenes (rc_scene_generate(scene_seed0 + s, n_boxes, n_emissive)). */
typedef struct {
    uint32_t version;
    uint32_t width, height, levels;  /* training G-buffer size and cascade count */
    uint64_t scene_seed0;
    uint32_t scenes, n_boxes, n_emissive;
    uint32_t epochs;
    uint32_t batch;                  /* samples per step, a multiple of 32 */
    float lr, lr_decay;              /* step size lr * lr_decay^epoch */
    uint32_t reserved;
    uint64_t seed;                   /* initial weights and shuffling */
} rc_train_config;
#define RC_TRAIN_MAX_EPOCHS 64
typedef struct {
    uint32_t version;
    uint32_t epochs;
    uint64_t samples;                /* training rows */
    uint64_t steps;
    float epoch_loss[RC_TRAIN_MAX_EPOCHS];
    double gen_ms, train_ms;         /* data generation and training wall time */
    float out_scale;
    uint32_t reserved;
} rc_train_report;
rc_status rc_surrogate_train(rc_ctx *ctx, const rc_train_config *cfg, rc_train_report *report);
/* Train on rows the caller made, e.g. with rc_surrogate_samples on G-buffers
 * read from dump files (rc_dump_read): feats [n_rows][RC_MLP_INPUTS] and
 * targets [n_rows][48] on the device, as rc_surrogate_samples writes them.
 * targets is scaled in place by 1 / out_scale.  Of cfg, only the optimiser
 * fields (epochs, batch, lr, lr_decay, seed) are used, and scenes is recorded
 * in the weights blob as the number of frames the rows came from.  Rows that
 * rc_surrogate_train would generate give the same network bit for bit.  No
 * counterpart in the reference (its renderer has no GI). */
rc_status rc_surrogate_train_rows(rc_ctx *ctx, const float *feats_dev, float *targets_dev, uint64_t n_rows,
                                  const rc_train_config *cfg, rc_train_report *report);
/* Readbacks of the last learned-far-field frame (device destinations, any may
 * be NULL): cone [Py_2][Px_2][16][4] (rgb, transmittance; needs a context
 * created with RC_FLAG_KEEP_DEBUG); level1 [Py_1][Px_1][4][4] merged level 1
 * (fp32); tile_err [ceil(Py_2/8)][ceil(Px_2/8)]. */
rc_status rc_debug_surrogate(rc_ctx *ctx, float *cone_dev, float *level1_dev, float *tile_err_dev, void *stream);
/* Mip level m >= 1 of the last learned-far-field frame: emission [h][w][4]
 * (mean E*[emitter] rgb, emitter area fraction) and coverage [h][w][4] (the
 * fractions of each cell's pixel columns and rows holding an emitter pixel,
 * then the same for dark occluder pixels). */
rc_status rc_debug_mip(rc_ctx *ctx, uint32_t level, float *emission_dev, float *coverage_dev, void *stream);

/* ---- batches of views (one graph launch) ------------------------------------
 * Independent views of uneven sizes (e.g. the sub-views of a renderer's
 * frame, or many cameras) run as ONE hipGraph launch.  No counterpart in the
 * reference, which renders one view per frame (src/window/app.rs:221-267).
 *
 * RC_BATCH_WORKLIST (default): a fixed handful of kernels per batch, each
 *   walking a work list of (view, stage block) jobs over ONE shared workspace:
 *   the exact distance transform's occupancy, column and row kernels, the
 *   split march of every (view, level >= 1), one launch per level index, and
 *   level 0 + gather: about 10 kernel nodes whatever the number of views.
 *   rc_batch_autotune tunes each stage for the batch as a whole.
 * RC_BATCH_STREAMS: one context per view; each view's frame is captured once
 *   and the views are dealt round-robin to `streams` captured streams so that
 *   small views run concurrently (about 17 kernel nodes per view).
 *   rc_batch_autotune tunes each distinct (width, height, levels) once.
 *
 * Either way every view's irradiance equals its single-view rc_run_device
 * frame (default distance transform, RC_SDF_EXACT) bit for bit.  The view buffers are the caller's and are bound at
 * create time (the graph keeps their addresses); depth is linear view depth.
 * Plain cascades only. */
typedef struct {
    uint32_t width, height;
    uint32_t levels;                 /* 0 = coverage rule */
    uint32_t reserved;
    const float *depth;              /* device [H][W] */
    const float *emissive;           /* device [3][H][W] */
    float *irradiance;               /* device [3][H][W] */
} rc_view;
#define RC_BATCH_WORKLIST 0u
#define RC_BATCH_STREAMS 1u
typedef struct {
    uint32_t storage;                /* RC_STORAGE_* */
    uint32_t mode;                   /* RC_BATCH_WORKLIST or RC_BATCH_STREAMS */
    uint32_t streams;                /* RC_BATCH_STREAMS: 1 .. 16 captured streams (ignored otherwise) */
    int device;
} rc_batch_config;
typedef struct rc_batch rc_batch;
rc_status rc_batch_create(const rc_view *views, uint32_t n_views, const rc_batch_config *cfg, rc_batch **out);
rc_status rc_batch_autotune(rc_batch *batch, uint32_t iters, void *stream);
rc_status rc_batch_run(rc_batch *batch, void *stream);
/* Kernel nodes of one batch launch (streams mode: known after the first
 * rc_batch_run) and the bytes of the shared workspace (work-list mode;
 * streams mode reports 0: each view's context owns its buffers). */
rc_status rc_batch_info(const rc_batch *batch, uint32_t *kernel_nodes, uint64_t *workspace_bytes);
/* levels of view i after the coverage rule */
rc_status rc_batch_view_levels(const rc_batch *batch, uint32_t i, uint32_t *levels);
const char *rc_batch_last_error(const rc_batch *batch);
void rc_batch_destroy(rc_batch *batch);

/* ---- cascade geometry (integer, identical to the kernels' indexing) ------- */
rc_status rc_probe_grid(uint32_t width, uint32_t height, uint32_t level, uint32_t *px, uint32_t *py);
uint64_t rc_level_rays(uint32_t width, uint32_t height, uint32_t level);
uint32_t rc_levels_for_coverage(uint32_t width, uint32_t height);

/* ---- seeded synthetic scenes (the benchmark's G-buffer producer) ---------- */
typedef struct {
    double lo[3], hi[3];
    double albedo[3];
    double emit[3];
} rc_box;
/* SplitMix64-seeded box scene in the unit-cube room (see the design notes'
 * scene specification).  boxes must hold n_boxes entries. */
rc_status rc_scene_generate(uint64_t seed, uint32_t n_boxes, uint32_t n_emissive, rc_box *boxes);
/* Analytic FP64 raster of the scene to a planar fp32 G-buffer on the device. */
rc_status rc_gbuffer_raster(const rc_box *boxes_host, uint32_t n_boxes, uint32_t width, uint32_t height,
                            float *depth_dev, float *albedo_dev, float *emissive_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
