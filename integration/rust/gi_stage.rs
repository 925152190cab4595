//! gi_stage.rs -- the radiance-cascade GI pass as a `RenderStage` of the
//! reference renderer (src/app.rs:3-7).  Copy into the renderer crate's
//! `src/`, add `mod gi_stage;` and the `rc-gi-sys` dependency; the stage runs
//! next to `DefaultRenderer` (src/renderer.rs:558), which keeps the raster.
//!
//! Call order is the reference's (src/window/app.rs:221-267): `update` ->
//! `render` -> submit.  `render` receives no queue (src/app.rs:4) and `update`
//! does (src/app.rs:6), so `update` uploads -- the G-buffer the raster left,
//! to the pass (`rc_update_host`), and the irradiance the last `render`
//! computed, to the texture the composite samples -- and `render` runs the
//! pass (`rc_render_host`).  The GI is one frame behind the raster, like the
//! reference's own per-frame uniform uploads (src/renderer.rs:620-631).
use std::ffi::CStr;
use std::os::raw::c_void;

use rc_gi_sys::*;

use crate::app::{AppState, RenderStage};

/// The renderer side of the stage: its raster targets read back into planar
/// fp32 planes (depth exactly as the Depth32Float target holds it,
/// src/texture.rs:149, which needs COPY_SRC, src/texture.rs:168; emissive rgb
/// from the materials' Ke), and the irradiance texture the composite samples.
pub trait GBufferIo {
    fn read_back(&mut self, state: &AppState, depth: &mut [f32], emissive: &mut [f32]);
    fn upload_irradiance(&mut self, queue: &wgpu::Queue, irradiance: &[f32]);
    fn resize(&mut self, device: &wgpu::Device, width: u32, height: u32);
}

/// Page-locked host plane (`rc_host_alloc`): copies to and from the GPU by
/// DMA instead of staged copies.
pub struct HostBuf {
    ptr: *mut f32,
    len: usize,
}

impl HostBuf {
    pub fn new(len: usize) -> Self {
        let mut p: *mut c_void = std::ptr::null_mut();
        unsafe { check(std::ptr::null(), rc_host_alloc((4 * len.max(1)) as u64, &mut p)) };
        HostBuf { ptr: p as *mut f32, len }
    }
    pub fn as_slice(&self) -> &[f32] {
        unsafe { std::slice::from_raw_parts(self.ptr, self.len) }
    }
    pub fn as_mut(&mut self) -> &mut [f32] {
        unsafe { std::slice::from_raw_parts_mut(self.ptr, self.len) }
    }
}

impl Drop for HostBuf {
    fn drop(&mut self) {
        unsafe { rc_host_free(self.ptr as *mut c_void) }
    }
}

/// The reference unwrap()s device failures (src/renderer.rs:176); the stage
/// does the same with the library's message.
fn check(ctx: *const rc_ctx, s: rc_status) {
    if s != RC_OK {
        let msg = unsafe { CStr::from_ptr(rc_last_error(ctx)) };
        panic!("rc_gi: status {s}: {}", msg.to_string_lossy());
    }
}

pub struct GiStage<IO: GBufferIo> {
    ctx: *mut rc_ctx,
    width: u32,
    height: u32,
    depth: HostBuf,
    emissive: HostBuf,
    irradiance: HostBuf,
    io: IO,
}

impl<IO: GBufferIo> GiStage<IO> {
    /// The camera of src/app.rs:26 (near 0.1, far 100) writes perspective_rh
    /// NDC depth (src/camera.rs:77-79); the pass converts it itself.
    pub fn new(width: u32, height: u32, io: IO) -> Self {
        let cfg = rc_config {
            version: RC_ABI_VERSION,
            width: width.max(1),
            height: height.max(1),
            levels: 0, // enough cascades to cover the diagonal
            storage: 0,
            flags: 0,
            device: 0,
            depth_mode: RC_DEPTH_NDC,
            z_near: 0.1,
            z_far: 100.0,
        };
        let mut ctx = std::ptr::null_mut();
        unsafe { check(std::ptr::null(), rc_create(&cfg, &mut ctx)) };
        let n = (cfg.width * cfg.height) as usize;
        GiStage {
            ctx,
            width: cfg.width,
            height: cfg.height,
            depth: HostBuf::new(n),
            emissive: HostBuf::new(3 * n),
            irradiance: HostBuf::new(3 * n),
            io,
        }
    }

    /// Stage times and counts of the frames so far (`rc_get_stats`).
    pub fn stats(&self) -> rc_stats {
        let mut s = std::mem::MaybeUninit::<rc_stats>::uninit();
        unsafe {
            check(self.ctx, rc_get_stats(self.ctx, s.as_mut_ptr()));
            s.assume_init()
        }
    }
}

impl<IO: GBufferIo> RenderStage<AppState> for GiStage<IO> {
    fn render(&self, _state: &mut AppState, _view: &wgpu::TextureView, _encoder: &mut wgpu::CommandEncoder) {
        // the stage owns the page-locked irradiance plane; render takes &self,
        // so the pass writes it through the raw pointer
        unsafe { check(self.ctx, rc_render_host(self.ctx, self.irradiance.ptr)) };
    }

    fn resize(&mut self, device: &wgpu::Device, config: &wgpu::SurfaceConfiguration) {
        self.width = config.width.max(1);
        self.height = config.height.max(1);
        unsafe { check(self.ctx, rc_resize(self.ctx, self.width, self.height, 0)) };
        let n = (self.width * self.height) as usize;
        self.depth = HostBuf::new(n);
        self.emissive = HostBuf::new(3 * n);
        self.irradiance = HostBuf::new(3 * n);
        self.io.resize(device, self.width, self.height);
    }

    fn update(&mut self, state: &AppState, queue: &wgpu::Queue) {
        self.io.upload_irradiance(queue, self.irradiance.as_slice());
        self.io.read_back(state, self.depth.as_mut(), self.emissive.as_mut());
        let g = rc_gbuffer { depth: self.depth.ptr, albedo: std::ptr::null(), emissive: self.emissive.ptr };
        unsafe { check(self.ctx, rc_update_host(self.ctx, &g)) };
    }
}

impl<IO: GBufferIo> Drop for GiStage<IO> {
    fn drop(&mut self) {
        unsafe { rc_destroy(self.ctx) }
    }
}
