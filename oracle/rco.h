/*
 * rco.h -- CPU oracle for the radiance-cascade GI path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a CPU restatement of the path's maths,
 * used as the checker by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py.  The product (radiancecascade_amd/, librcgi.so) never
 * links, loads or calls it.
 *
 * The reference repository (jw910731/RadianceCascade) contains no
 * radiance-cascade code: its only pass is the Blinn-Phong forward raster
 * (src/renderer.rs:559-613, src/shader.wgsl:76-101).  The maths restated here
 * therefore follows the published method -- Sannikov, "Radiance Cascades"
 * (2023) for the cascade hierarchy and bilinear merge; for the distance field
 * the exact Euclidean distance transform (Felzenszwalb & Huttenlocher,
 * "Distance Transforms of Sampled Functions", Theory of Computing 8, 2012),
 * or on request Rong & Tan, "Jump Flooding in GPU" (I3D 2006) -- with every
 * free choice
 * frozen in the specification section of the design notes.  Where the path
 * touches the reference (depth convention, pass boundary, albedo) the function
 * cites the reference file:line it follows.
 *
 * Two arithmetic flavours of the cascade stages exist:
 *   *_f64  the FP64 oracle (the error judge for every floating-point output);
 *   *_f32  an FP32 restatement whose operation order (explicit fmaf, the
 *          bilinear and 4-child reduction order) is the one the HIP kernels
 *          use, so it is also the CPU baseline.
 * Integer stages (scene PRNG, occupancy, jump flood, probe/ray indexing, hit
 * pixel indices) have one flavour and are compared bit-exactly.
 *
 * Parity pins (see tests/test_oracle_*.py): SplitMix64 published vector;
 * raster closed forms (room-only depth); exact transform == scipy's exact EDT
 * and the brute-force nearest seed on fields of every density and shape;
 * JFA == brute-force nearest seed for a single seed and within the JFA error
 * bound otherwise; empty-scene and
 * constant-upper-level closed forms of march/merge/gather; analytic
 * emissive-disc hit test of the march; FP32 vs FP64 agreement.
 */
#ifndef RCO_H
#define RCO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RCO_MAX_BOXES 256
#define RCO_MAX_LEVELS 8
#define RCO_SEED_NONE 0xFFFFFFFFu

/* Hit codes returned per ray by the march (>= 0 is the hit pixel y*W+x). */
#define RCO_RAY_EXIT_SCREEN (-1)
#define RCO_RAY_REACHED_T1  (-2)

typedef struct {
    double lo[3], hi[3];
    double albedo[3];
    double emit[3];
} rco_box;

/* ---- scene and G-buffer (H0) ---------------------------------------- */
uint64_t rco_splitmix64_next(uint64_t *state);
double   rco_u01(uint64_t *state);
int      rco_scene_generate(uint64_t seed, int n_boxes, int n_emissive, rco_box *boxes);
void     rco_raster(const rco_box *boxes, int n, int W, int H,
                    float *depth, float *albedo, float *emissive);

/* ---- cascade geometry (H3) ------------------------------------------ */
int      rco_probe_spacing(int level);             /* s_i = 2 * 2^i      */
int      rco_num_dirs(int level);                  /* N_i = 4 * 4^i      */
void     rco_probe_grid(int W, int H, int level, int *px, int *py);
int64_t  rco_interval_t0(int level);               /* 2(4^i - 1)/3       */
int64_t  rco_interval_t1(int level);               /* 2(4^(i+1) - 1)/3   */
int64_t  rco_level_rays(int W, int H, int level);  /* R_i                */
int      rco_levels_for_coverage(int W, int H);
int64_t  rco_ray_index(int W, int H, int level, int px, int py, int k);
void     rco_dir_table_f32(int level, float *dirs /* [N_i][2] */);

/* ---- occupancy + jump flood (H1, H2) -------------------------------- */
int      rco_jfa_num_passes(int W, int H);
void     rco_seed_init(int W, int H, const float *depth, const float *emissive,
                       uint32_t *seed);
void     rco_jfa_pass(int W, int H, int step, const uint32_t *in, uint32_t *out);
void     rco_jfa(int W, int H, uint32_t *seed, uint32_t *tmp);
double   rco_sdf_value_d2(uint64_t d2);
void     rco_sdf_f32(int W, int H, const uint32_t *seed, float *sdf);
void     rco_sdf_f64(int W, int H, const uint32_t *seed, double *sdf);
void     rco_nearest_seed_bruteforce(int W, int H, const uint32_t *seed_init,
                                     uint32_t *nearest);
/* exact squared distance to the nearest occupied pixel (-1: none occupied),
 * and the field the march reads from it                                  */
void     rco_edt_d2(int W, int H, const uint32_t *seed_init, int64_t *d2);
void     rco_sdf_from_d2_f32(int W, int H, const int64_t *d2, float *sdf);
void     rco_sdf_from_d2_f64(int W, int H, const int64_t *d2, double *sdf);
/* ground-truth irradiance by an exact per-pixel grid walk (rco.c) */
void     rco_bruteforce_irradiance(int W, int H, const uint32_t *seed_init, const float *emissive, int n_dirs,
                                   double t_max, const int32_t *pixels, int n_pix, double *out);

/* ---- march / merge / gather (H4-H6), both flavours ------------------- */
/* rgba: [Py][Px][N][4] (canonical ray order), hit: [Py][Px][N] or NULL */
void rco_march_level_f64(int W, int H, int level, const double *sdf,
                         const float *emissive, double *rgba, int32_t *hit);
void rco_march_level_f32(int W, int H, int level, const float *sdf,
                         const float *emissive, float *rgba, int32_t *hit);
/* A_i[Py][Px][N_i/4][4]: merged level `level`, averaged over the 4 children of
 * each level-(i-1) direction.  upper = A_{i+1} or NULL for the top level.    */
void rco_merge_level_f64(int W, int H, int level, const double *march_rgba,
                         const double *upper, double *A_out);
void rco_merge_level_f32(int W, int H, int level, const float *march_rgba,
                         const float *upper, float *A_out);
/* irradiance [3][H][W] from A_0 ([Py0][Px0][1][4]) */
void rco_gather_f64(int W, int H, const double *A0, double *irr);
void rco_gather_f32(int W, int H, const float *A0, float *irr);

/* Whole path from the G-buffer.  levels_out (optional, may be NULL): array of
 * L pointers, entry i receives A_i (size Py_i*Px_i*max(N_i/4,1)*4).
 * sdf_mode: RCO_SDF_EXACT (the distance field from the exact transform) or
 * RCO_SDF_JFA (from the jump flood).  seeds_out (optional, RCO_SDF_JFA only)
 * receives the final JFA field.                                             */
#define RCO_SDF_EXACT 0
#define RCO_SDF_JFA   1
int rco_run_f64(int W, int H, int L, const float *depth, const float *emissive,
                double *irr, double **levels_out, uint32_t *seeds_out, int sdf_mode);
int rco_run_f32(int W, int H, int L, const float *depth, const float *emissive,
                float *irr, float **levels_out, uint32_t *seeds_out, int sdf_mode);

/* OpenMP thread count for the f32/f64 level loops (cpu_baseline). */
void rco_set_threads(int n);
int  rco_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
