//! Raw bindings to `librcgi.so`, one declaration per entry point and struct of
//! `include/rc_gi.h` (ABI version 2), in the header's order.  Safe wrappers
//! and the renderer stage live in the renderer crate (`../gi_stage.rs`).
//! Every function returns an `rc_status` (0 = OK) unless stated; on failure
//! `rc_last_error` (or the batch / dump / noalloc variants) gives the message.
#![allow(non_camel_case_types)]
use std::os::raw::{c_char, c_int, c_void};

pub const RC_ABI_VERSION: u32 = 2;
pub const RC_MAX_LEVELS: usize = 8;
pub const RC_MAX_DIM: u32 = 16384;
pub const RC_MAX_BOXES: u32 = 256;
pub const RC_FLAG_PROFILE: u32 = 0x1;
pub const RC_FLAG_KEEP_DEBUG: u32 = 0x2;
pub const RC_FLAG_GRAPH: u32 = 0x8;
pub const RC_DEPTH_LINEAR: u32 = 0;
pub const RC_DEPTH_NDC: u32 = 1;
pub const RC_SDF_EXACT: u32 = 0;
pub const RC_SDF_JUMP_FLOOD: u32 = 1;
pub const RC_STAGE_SEED: usize = 0;
pub const RC_STAGE_JFA: usize = 1;
pub const RC_STAGE_JFA_TAIL: usize = 2;
pub const RC_STAGE_LEVEL0: usize = 3;
pub const RC_STAGE_MARCH: usize = RC_STAGE_LEVEL0 + RC_MAX_LEVELS;
pub const RC_STAGE_NEAR_SDF: usize = RC_STAGE_MARCH + 1;
pub const RC_STAGE_MIP: usize = RC_STAGE_MARCH + 2;
pub const RC_STAGE_CONE: usize = RC_STAGE_MARCH + 3;
pub const RC_STAGE_SURROGATE: usize = RC_STAGE_MARCH + 4;
pub const RC_STAGE_FALLBACK: usize = RC_STAGE_MARCH + 5;
pub const RC_NUM_STAGES: usize = RC_STAGE_MARCH + 6;
pub const RC_NUM_TILE_SHAPES: u32 = 5;
pub const RC_MLP_INPUTS: usize = 64;
pub const RC_MLP_OUTPUTS: usize = 64;
pub const RC_MLP_PARAMS: usize = 16640;
pub const RC_TRAIN_MAX_EPOCHS: usize = 64;
pub const RC_BATCH_WORKLIST: u32 = 0;
pub const RC_BATCH_STREAMS: u32 = 1;

pub type rc_status = i32;
pub const RC_OK: rc_status = 0;

#[repr(C)]
pub struct rc_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct rc_batch {
    _p: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rc_config {
    pub version: u32,
    pub width: u32,
    pub height: u32,
    pub levels: u32,
    pub storage: u32,
    pub flags: u32,
    pub device: i32,
    pub depth_mode: u32,
    pub z_near: f32,
    pub z_far: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rc_gbuffer {
    pub depth: *const f32,
    pub albedo: *const f32,
    pub emissive: *const f32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct rc_stats {
    pub version: u32,
    pub runs: u32,
    pub stage_ms: [f64; RC_NUM_STAGES],
    pub stage_launches: [u32; RC_NUM_STAGES],
    pub level_rays: [u64; RC_MAX_LEVELS],
    pub level_bytes_stored: [u64; RC_MAX_LEVELS],
    pub surrogate_frames: u64,
    pub fallback_frames: u64,
    pub flagged_tiles: u64,
    pub tiles_per_frame: u64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct rc_plane {
    pub name: [c_char; 16],
    pub channels: u32,
    pub height: u32,
    pub width: u32,
    pub reserved: u32,
    pub data: *const f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rc_surrogate_config {
    pub version: u32,
    pub mode: u32,
    pub threshold: f32,
    pub reserved: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rc_train_config {
    pub version: u32,
    pub width: u32,
    pub height: u32,
    pub levels: u32,
    pub scene_seed0: u64,
    pub scenes: u32,
    pub n_boxes: u32,
    pub n_emissive: u32,
    pub epochs: u32,
    pub batch: u32,
    pub lr: f32,
    pub lr_decay: f32,
    pub reserved: u32,
    pub seed: u64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct rc_train_report {
    pub version: u32,
    pub epochs: u32,
    pub samples: u64,
    pub steps: u64,
    pub epoch_loss: [f32; RC_TRAIN_MAX_EPOCHS],
    pub gen_ms: f64,
    pub train_ms: f64,
    pub out_scale: f32,
    pub reserved: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rc_view {
    pub width: u32,
    pub height: u32,
    pub levels: u32,
    pub reserved: u32,
    pub depth: *const f32,
    pub emissive: *const f32,
    pub irradiance: *mut f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rc_batch_config {
    pub storage: u32,
    pub mode: u32,
    pub streams: u32,
    pub device: c_int,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct rc_box {
    pub lo: [f64; 3],
    pub hi: [f64; 3],
    pub albedo: [f64; 3],
    pub emit: [f64; 3],
}

#[link(name = "rcgi")]
extern "C" {
    // the pass (RenderStage, src/app.rs:3-7)
    pub fn rc_create(cfg: *const rc_config, out: *mut *mut rc_ctx) -> rc_status;
    pub fn rc_resize(ctx: *mut rc_ctx, width: u32, height: u32, levels: u32) -> rc_status;
    pub fn rc_update_host(ctx: *mut rc_ctx, host: *const rc_gbuffer) -> rc_status;
    pub fn rc_render_host(ctx: *mut rc_ctx, irradiance_host: *mut f32) -> rc_status;
    pub fn rc_host_alloc(bytes: u64, out: *mut *mut c_void) -> rc_status;
    pub fn rc_host_free(ptr: *mut c_void);
    pub fn rc_run_device(ctx: *mut rc_ctx, dev: *const rc_gbuffer, irradiance_dev: *mut f32, stream: *mut c_void)
        -> rc_status;
    pub fn rc_run_device_noalloc(width: u32, height: u32, levels: u32, dev: *const rc_gbuffer,
                                 irradiance_dev: *mut f32, stream: *mut c_void) -> rc_status;
    pub fn rc_noalloc_scratch_bytes(width: u32, height: u32, levels: u32) -> u64;
    pub fn rc_noalloc_last_error() -> *const c_char;
    pub fn rc_destroy(ctx: *mut rc_ctx);

    // tuning
    pub fn rc_set_level_tiling(ctx: *mut rc_ctx, level: u32, shape: u32, rays_per_lane: u32) -> rc_status;
    pub fn rc_get_level_tiling(ctx: *const rc_ctx, level: u32, shape: *mut u32, rays_per_lane: *mut u32)
        -> rc_status;
    pub fn rc_autotune(ctx: *mut rc_ctx, dev: *const rc_gbuffer, irradiance_dev: *mut f32, iters: u32,
                       stream: *mut c_void) -> rc_status;
    pub fn rc_set_sdf_mode(ctx: *mut rc_ctx, mode: u32) -> rc_status;
    pub fn rc_get_sdf_mode(ctx: *const rc_ctx) -> u32;
    pub fn rc_set_jfa_fuse(ctx: *mut rc_ctx, fused_step: u32) -> rc_status;
    pub fn rc_get_jfa_fuse(ctx: *const rc_ctx) -> u32;
    pub fn rc_set_jfa_rows(ctx: *mut rc_ctx, rows: u32) -> rc_status;
    pub fn rc_get_jfa_rows(ctx: *const rc_ctx) -> u32;
    pub fn rc_set_march(ctx: *mut rc_ctx, split: u32, speculation: u32, rays_per_lane: u32) -> rc_status;
    pub fn rc_get_march(ctx: *const rc_ctx, split: *mut u32, speculation: *mut u32, rays_per_lane: *mut u32)
        -> rc_status;
    pub fn rc_set_march_order(ctx: *mut rc_ctx, level: u32, order: u32) -> rc_status;
    pub fn rc_get_march_order(ctx: *const rc_ctx, level: u32, order: *mut u32) -> rc_status;

    // state, errors, devices
    pub fn rc_abi_version() -> u32; // compare with RC_ABI_VERSION before rc_create
    pub fn rc_get_config(ctx: *const rc_ctx, out: *mut rc_config) -> rc_status;
    pub fn rc_get_stats(ctx: *const rc_ctx, out: *mut rc_stats) -> rc_status;
    pub fn rc_reset_stats(ctx: *mut rc_ctx) -> rc_status;
    pub fn rc_last_error(ctx: *const rc_ctx) -> *const c_char;
    pub fn rc_status_string(s: rc_status) -> *const c_char;
    pub fn rc_device_count(count: *mut c_int) -> rc_status;

    // parity readback
    pub fn rc_debug_seeds(ctx: *mut rc_ctx, dev_out: *mut u32, stream: *mut c_void) -> rc_status;
    pub fn rc_debug_sdf(ctx: *mut rc_ctx, dev_out: *mut f32, stream: *mut c_void) -> rc_status;
    pub fn rc_debug_level(ctx: *mut rc_ctx, level: u32, dev_out: *mut f32, stream: *mut c_void) -> rc_status;
    pub fn rc_debug_march(ctx: *mut rc_ctx, level: u32, hit_dev: *mut i32, index_dev: *mut i64,
                          stream: *mut c_void) -> rc_status;
    pub fn rc_error_norms(ctx: *mut rc_ctx, x: *const f32, reference: *const c_void, ref_is_f64: u32, n: u64,
                          out_dev: *mut f64, stream: *mut c_void) -> rc_status;

    // presented frame (replaces DefaultRenderer::render's output, src/renderer.rs:559-613)
    pub fn rc_composite(ctx: *mut rc_ctx, albedo: *const f32, emissive: *const f32, irradiance: *const f32,
                        direct: *const f32, bgra8: *mut u32, stream: *mut c_void) -> rc_status;
    pub fn rc_srgb_thresholds(t: *mut f32);

    // dump files (RCDUMP01)
    pub fn rc_dump_write(path: *const c_char, planes: *const rc_plane, n_planes: u32) -> rc_status;
    pub fn rc_dump_info(path: *const c_char, planes: *mut rc_plane, cap: u32, n_out: *mut u32) -> rc_status;
    pub fn rc_dump_read(path: *const c_char, name: *const c_char, dst: *mut f32, count: u64) -> rc_status;
    pub fn rc_dump_last_error() -> *const c_char;
    pub fn rc_sha256(data: *const c_void, n: u64, out32: *mut u8);

    // learned far field
    pub fn rc_set_surrogate(ctx: *mut rc_ctx, cfg: *const rc_surrogate_config) -> rc_status;
    pub fn rc_get_surrogate(ctx: *const rc_ctx, out: *mut rc_surrogate_config) -> rc_status;
    pub fn rc_mlp_set(ctx: *mut rc_ctx, params: *const f32, in_mean: *const f32, in_istd: *const f32,
                      out_scale: f32) -> rc_status;
    pub fn rc_mlp_get(ctx: *const rc_ctx, params: *mut f32, in_mean: *mut f32, in_istd: *mut f32,
                      out_scale: *mut f32) -> rc_status;
    pub fn rc_surrogate_save(ctx: *const rc_ctx, buf: *mut c_void, capacity: u64, size: *mut u64) -> rc_status;
    pub fn rc_surrogate_load(ctx: *mut rc_ctx, buf: *const c_void, size: u64) -> rc_status;
    pub fn rc_mlp_forward(ctx: *mut rc_ctx, feats_dev: *const f32, n: u32, out_dev: *mut f32, stream: *mut c_void)
        -> rc_status;
    pub fn rc_mlp_gradient(ctx: *mut rc_ctx, feats_dev: *const f32, targets_dev: *const f32, n: u32,
                           grad_dev: *mut f32, loss_dev: *mut f32, stream: *mut c_void) -> rc_status;
    pub fn rc_mlp_adam(ctx: *mut rc_ctx, grad_dev: *const f32, lr: f32, stream: *mut c_void) -> rc_status;
    pub fn rc_surrogate_samples(ctx: *mut rc_ctx, dev: *const rc_gbuffer, feats_dev: *mut f32,
                                targets_dev: *mut f32, stream: *mut c_void) -> rc_status;
    pub fn rc_surrogate_train(ctx: *mut rc_ctx, cfg: *const rc_train_config, report: *mut rc_train_report)
        -> rc_status;
    pub fn rc_surrogate_train_rows(ctx: *mut rc_ctx, feats_dev: *const f32, targets_dev: *mut f32, n_rows: u64,
                                   cfg: *const rc_train_config, report: *mut rc_train_report) -> rc_status;
    pub fn rc_debug_surrogate(ctx: *mut rc_ctx, cone_dev: *mut f32, level1_dev: *mut f32, tile_err_dev: *mut f32,
                              stream: *mut c_void) -> rc_status;
    pub fn rc_debug_mip(ctx: *mut rc_ctx, level: u32, emission_dev: *mut f32, coverage_dev: *mut f32,
                        stream: *mut c_void) -> rc_status;

    // batches of views
    pub fn rc_batch_create(views: *const rc_view, n_views: u32, cfg: *const rc_batch_config,
                           out: *mut *mut rc_batch) -> rc_status;
    pub fn rc_batch_autotune(batch: *mut rc_batch, iters: u32, stream: *mut c_void) -> rc_status;
    pub fn rc_batch_run(batch: *mut rc_batch, stream: *mut c_void) -> rc_status;
    pub fn rc_batch_info(batch: *const rc_batch, kernel_nodes: *mut u32, workspace_bytes: *mut u64) -> rc_status;
    pub fn rc_batch_view_levels(batch: *const rc_batch, i: u32, levels: *mut u32) -> rc_status;
    pub fn rc_batch_last_error(batch: *const rc_batch) -> *const c_char;
    pub fn rc_batch_destroy(batch: *mut rc_batch);

    // cascade geometry and seeded scenes
    pub fn rc_probe_grid(width: u32, height: u32, level: u32, px: *mut u32, py: *mut u32) -> rc_status;
    pub fn rc_level_rays(width: u32, height: u32, level: u32) -> u64;
    pub fn rc_levels_for_coverage(width: u32, height: u32) -> u32;
    pub fn rc_scene_generate(seed: u64, n_boxes: u32, n_emissive: u32, boxes: *mut rc_box) -> rc_status;
    pub fn rc_gbuffer_raster(boxes_host: *const rc_box, n_boxes: u32, width: u32, height: u32,
                             depth_dev: *mut f32, albedo_dev: *mut f32, emissive_dev: *mut f32,
                             stream: *mut c_void) -> rc_status;
}
