"""MI355X-native radiance-cascade GI pass (HIP kernels for gfx950 behind a C ABI).

The pass replaces the reference renderer's ``RenderStage`` slot
(src/app.rs:3-7) with jump flood -> per-level interval march and top-down
merge -> level-0 gather to per-pixel irradiance.  The kernels live in
``librcgi.so`` (built from ``csrc/`` by ``radiancecascade_amd.build``); there
is no CPU fallback and importing the stage fails if the library is absent.
"""
from ._abi import RcError, load  # noqa: F401
from .stage import (RadianceCascadeBatch, RadianceCascadeStage, batch_views_spec, device_count, dump_info, dump_read,  # noqa: F401
                    dump_write, gbuffer_raster, host_buffer, level_rays, levels_for_coverage, packaged_weights, probe_grid,
                    scene_generate, sha256, srgb_thresholds)

__all__ = ["RadianceCascadeBatch", "RadianceCascadeStage", "batch_views_spec", "RcError", "device_count", "gbuffer_raster", "level_rays",
           "levels_for_coverage", "load", "packaged_weights", "probe_grid", "scene_generate", "srgb_thresholds",
           "dump_write", "dump_info", "dump_read", "sha256", "host_buffer"]
