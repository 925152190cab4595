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

// ---- occupancy + jump flood (rc_jfa.hip) --------------------------------------
int jfa_num_passes(int W, int H);
// largest step fused into the LDS tail for a requested fuse limit (0 = none)
int jfa_tail_step(int W, int H, int fuse);
hipError_t launch_sdf_untile(const float *sdf, float *out, int W, int H, hipStream_t s);
hipError_t launch_seed_init(const float *depth, const float *emissive, uint32_t *seed, int W, int H, hipStream_t s);
hipError_t launch_jfa_pass(const uint32_t *in, uint32_t *out, float *sdf, int W, int H, int step, bool final_pass,
                           hipStream_t s);
hipError_t launch_jfa_tail(const uint32_t *in, uint32_t *seeds_out, float *sdf, int W, int H, int tail_step,
                           hipStream_t s);

// ---- cascades (rc_cascade.hip) ----------------------------------------------
// Workgroup tile of a level kernel: TX x TY probes x KD directions, TX*TY*KD = 256.
enum TileShape : int { TILE_8x8x4 = 0, TILE_4x4x16 = 1, TILE_2x2x64 = 2, TILE_1x1x256 = 3, TILE_COUNT = 4 };
int tile_kd(int shape);
int tile_tx(int shape);
int tile_ty(int shape);

struct LevelArgs {
    const float *sdf;        // [H][W]
    const float *emissive;   // [3][H][W]
    const float2 *dirs;      // [N] unit directions of this level
    const void *upper;       // merged level i+1 (storage texels), nullptr at the top level
    void *out;               // merged level i (storage texels)
    int W, H;
    int level, PX, PY, N;    // this level
    int UX, UY;              // probe grid of level i+1
    float t0, t1, spacing;
    const int *codes;        // pre-marched ray codes of this level, or nullptr (march in place)
};

// k_march_all: every level >= 1 in one launch (level-major, top level first).
struct MarchLevel {
    const float2 *dirs;
    int *codes;              // [PY][PX][N] ray codes
    int PX, PY, N, log2n;
    float inv_px;            // 1/PX in fp32 (probe -> row, corrected)
    float t0, t1, spacing;
    int block_begin, blocks; // workgroups of 256*M canonical rays
};
struct MarchAllArgs {
    const float *sdf;
    int W, H, nlev;
    MarchLevel lev[RC_MAX_LEVELS];
};
hipError_t launch_march_all(const MarchAllArgs &a, int speculation, int rays_per_lane, hipStream_t s);

hipError_t launch_level(const LevelArgs &a, Storage st, int shape, int rays_per_lane, hipStream_t s);
hipError_t launch_level0_gather(const LevelArgs &a, Storage st, float *irradiance, float *debug_fluence,
                                hipStream_t s);
hipError_t launch_level_to_f32(const void *level, Storage st, size_t texels, float *out, hipStream_t s);
hipError_t launch_march_probe(const LevelArgs &a, int spec, int32_t *hit, int64_t *index, hipStream_t s);

// ---- error norms (rc_norms.hip) ---------------------------------------------
size_t norm_workspace_bytes();
hipError_t launch_error_norms(const float *x, const void *ref, bool ref_f64, size_t n, double *workspace,
                              double *out, hipStream_t s);

}  // namespace rc
