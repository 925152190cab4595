/* The oracle's whole path under AddressSanitizer / UndefinedBehaviorSanitizer
 * (host only): scenes, raster, jump flood, exact transform, both flavours of the cascades with
 * every level and the seed field read back, on ragged and degenerate shapes.
 * Exits non-zero on a failed invariant; the sanitizers abort on any memory
 * or UB error. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/rco.h"

static int check_shape(int W, int H, int L, uint64_t seed)
{
    rco_box boxes[16];
    const int nb = rco_scene_generate(seed, 12, 3, boxes);
    if (nb != 12) return fprintf(stderr, "scene count %d\n", nb), 1;
    const size_t n = (size_t)W * H;
    float *depth = malloc(n * sizeof(float)), *alb = malloc(3 * n * sizeof(float)), *em = malloc(3 * n * sizeof(float));
    rco_raster(boxes, nb, W, H, depth, alb, em);
    double *irr64 = malloc(3 * n * sizeof(double));
    float *irr32 = malloc(3 * n * sizeof(float));
    uint32_t *seeds = malloc(n * sizeof(uint32_t));
    double *lv64[12];
    float *lv32[12];
    for (int i = 0; i < L; ++i) {
        int px, py;
        rco_probe_grid(W, H, i, &px, &py);
        const int k = rco_num_dirs(i) / 4 > 1 ? rco_num_dirs(i) / 4 : 1;
        lv64[i] = malloc((size_t)px * py * k * 4 * sizeof(double));
        lv32[i] = malloc((size_t)px * py * k * 4 * sizeof(float));
    }
    int bad = rco_run_f64(W, H, L, depth, em, irr64, lv64, seeds, RCO_SDF_JFA) != 0;
    bad |= rco_run_f32(W, H, L, depth, em, irr32, lv32, NULL, RCO_SDF_JFA) != 0;
    for (size_t i = 0; i < 3 * n && !bad; ++i)
        if (!isfinite(irr64[i]) || irr64[i] < 0 || fabs(irr64[i] - irr32[i]) > 1e-2 * (1 + fabs(irr64[i]))) bad = 1;
    /* the jump flood never finds a nearer seed than the exact search */
    uint32_t *init = malloc(n * sizeof(uint32_t)), *near = malloc(n * sizeof(uint32_t));
    rco_seed_init(W, H, depth, em, init);
    rco_nearest_seed_bruteforce(W, H, init, near);
    for (size_t p = 0; p < n && !bad; ++p) {
        if ((near[p] == 0xFFFFFFFFu) != (seeds[p] == 0xFFFFFFFFu)) bad = 1;
        if (near[p] == 0xFFFFFFFFu) continue;
        const int x = (int)(p % W), y = (int)(p / W);
        const int64_t nx = (int64_t)(near[p] & 0xFFFF) - x, ny = (int64_t)(near[p] >> 16) - y;
        const int64_t jx = (int64_t)(seeds[p] & 0xFFFF) - x, jy = (int64_t)(seeds[p] >> 16) - y;
        const int64_t dn = nx * nx + ny * ny, dj = jx * jx + jy * jy;
        if (dj < dn) bad = 1;
    }
    /* the exact transform equals the brute-force nearest distance, and the
     * exact-field frame is finite */
    int64_t *d2 = malloc(n * sizeof(int64_t));
    rco_edt_d2(W, H, init, d2);
    for (size_t p = 0; p < n && !bad; ++p) {
        if (near[p] == 0xFFFFFFFFu) { if (d2[p] != -1) bad = 1; continue; }
        const int x = (int)(p % W), y = (int)(p / W);
        const int64_t nx = (int64_t)(near[p] & 0xFFFF) - x, ny = (int64_t)(near[p] >> 16) - y;
        if (d2[p] != nx * nx + ny * ny) bad = 1;
    }
    bad |= rco_run_f32(W, H, L, depth, em, irr32, lv32, NULL, RCO_SDF_EXACT) != 0;
    for (size_t i = 0; i < 3 * n && !bad; ++i)
        if (!isfinite(irr32[i]) || irr32[i] < 0) bad = 1;
    free(d2);
    for (int i = 0; i < L; ++i) {
        free(lv64[i]);
        free(lv32[i]);
    }
    free(depth); free(alb); free(em); free(irr64); free(irr32); free(seeds); free(init); free(near);
    if (bad) fprintf(stderr, "shape %dx%d L=%d failed\n", W, H, L);
    return bad;
}

int main(void)
{
    rco_set_threads(1);
    const int shapes[][3] = {{1, 1, 1}, {3, 2, 2}, {2, 3, 2}, {300, 1, 5}, {1, 97, 4}, {97, 61, 0}, {128, 96, 5}, {64, 64, 4}};
    int bad = 0;
    for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); ++s) {
        const int W = shapes[s][0], H = shapes[s][1];
        const int L = shapes[s][2] ? shapes[s][2] : rco_levels_for_coverage(W, H);
        bad |= check_shape(W, H, L, 7 + s);
    }
    printf(bad ? "oracle driver: FAILED\n" : "oracle driver: ok\n");
    return bad;
}
