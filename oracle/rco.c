/*
 * rco.c -- CPU oracle: scene generator, G-buffer raster, cascade geometry,
 * occupancy, exact distance transform and jump flood.  TEST INFRASTRUCTURE
 * ONLY (see rco.h): only the
 * tests, the smoke check and bench.py's cpu_baseline leg load it.
 *
 * Built with -ffp-contract=off so every float/double operation is the single
 * IEEE-rounded operation written here (the G-buffer raster and the occupancy
 * test must round exactly as the GPU does for the jump flood to match
 * bit-exactly).
 */
#include "rco.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* PRNG: SplitMix64 (Steele, Lea, Flood, OOPSLA 2014; Vigna's public-domain  */
/* splitmix64.c).  Seed 0 yields e220a8397b1dcdaf, 6e789e6aa1b965f4, ...     */
/* The reference has no PRNG; the scene is synthetic (no reference anchor).  */
/* ------------------------------------------------------------------------ */
uint64_t rco_splitmix64_next(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double rco_u01(uint64_t *state)
{
    return (double)(rco_splitmix64_next(state) >> 11) * 0x1.0p-53;
}

/*
 * Seeded box scene inside the unit-cube room.  Per box, 9 draws in the order
 * centre x,y,z; edge x,y,z in [0.1,0.4]; albedo r,g,b in [0.2,0.9].  A box
 * that overlaps the open clear zone (0.35,0.65)^3 around the camera at
 * (0.5,0.5,0.5) is redrawn (all 9 draws): without it a box a hair in front of
 * the lens can fill the whole frame and leave no distance field.  After all
 * boxes, the first n_emissive boxes draw an emitted radiance r,g,b in [1,10].
 * No reference anchor: the reference loads OBJ scenes (src/primitives.rs:122)
 * and has no emissive material field (src/primitives.rs:75-83).
 */
int rco_scene_generate(uint64_t seed, int n_boxes, int n_emissive, rco_box *boxes)
{
    if (n_boxes < 0 || n_boxes > RCO_MAX_BOXES || n_emissive < 0 || n_emissive > n_boxes)
        return -1;
    uint64_t st = seed;
    for (int b = 0; b < n_boxes; ++b) {
        for (;;) {
            double c[3], e[3], a[3];
            for (int j = 0; j < 3; ++j) c[j] = rco_u01(&st);
            for (int j = 0; j < 3; ++j) e[j] = 0.1 + 0.3 * rco_u01(&st);
            for (int j = 0; j < 3; ++j) a[j] = 0.2 + 0.7 * rco_u01(&st);
            int inside = 1;
            for (int j = 0; j < 3; ++j) {
                double lo = c[j] - 0.5 * e[j], hi = c[j] + 0.5 * e[j];
                boxes[b].lo[j] = lo;
                boxes[b].hi[j] = hi;
                boxes[b].albedo[j] = a[j];
                boxes[b].emit[j] = 0.0;
                if (!(lo < 0.65 && 0.35 < hi)) inside = 0;
            }
            if (!inside) break;
        }
    }
    for (int b = 0; b < n_emissive; ++b)
        for (int j = 0; j < 3; ++j) boxes[b].emit[j] = 1.0 + 9.0 * rco_u01(&st);
    return n_boxes;
}

/*
 * Analytic G-buffer.  Pinhole camera at (0.5,0.5,0.5), forward +z, up +y,
 * right = forward x up = -x: the look_to_rh frame of src/camera.rs:43-52.
 * Vertical field of view 60 degrees (tan 30 = 1/sqrt 3) as the fovy of
 * perspective_rh (src/camera.rs:77-79).  Depth is the linear view-space
 * distance along the optical axis (the reference's NDC depth d maps to it by
 * D = n f / (f - d (f - n))).  Planar fp32 outputs: depth [H][W],
 * albedo [3][H][W], emissive [3][H][W].  Room albedo 0.75, not emissive.
 * Nearest hit wins; on equal distance the room, then the lower box index.
 */
void rco_raster(const rco_box *boxes, int n, int W, int H,
                float *depth, float *albedo, float *emissive)
{
    const double th = 1.0 / sqrt(3.0);
    const double aspect = (double)W / (double)H;
    const double o = 0.5;
    const size_t plane = (size_t)W * (size_t)H;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            double ndc_x = (2.0 * ((double)x + 0.5)) / (double)W - 1.0;
            double ndc_y = 1.0 - (2.0 * ((double)y + 0.5)) / (double)H;
            double d[3];
            d[0] = -((ndc_x * th) * aspect);
            d[1] = ndc_y * th;
            d[2] = 1.0;
            double best = INFINITY;
            for (int a = 0; a < 3; ++a) {
                double t;
                if (d[a] > 0.0) t = (1.0 - o) / d[a];
                else if (d[a] < 0.0) t = (0.0 - o) / d[a];
                else continue;
                if (t < best) best = t;
            }
            int obj = -1;
            for (int b = 0; b < n; ++b) {
                double tn = -INFINITY, tf = INFINITY;
                for (int a = 0; a < 3; ++a) {
                    double t1 = (boxes[b].lo[a] - o) / d[a];
                    double t2 = (boxes[b].hi[a] - o) / d[a];
                    if (t1 > t2) { double s = t1; t1 = t2; t2 = s; }
                    if (t1 > tn) tn = t1;
                    if (t2 < tf) tf = t2;
                }
                if (tn <= tf && tn > 0.0 && tn < best) { best = tn; obj = b; }
            }
            size_t p = (size_t)y * (size_t)W + (size_t)x;
            depth[p] = (float)(best * d[2]);
            for (int c = 0; c < 3; ++c) {
                albedo[c * plane + p] = obj < 0 ? 0.75f : (float)boxes[obj].albedo[c];
                emissive[c * plane + p] = obj < 0 ? 0.0f : (float)boxes[obj].emit[c];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Cascade geometry.  Level i: spacing s_i = 2*2^i px, N_i = 4*4^i           */
/* directions, ceil-sized probe grid, interval [2(4^i-1)/3, 2(4^(i+1)-1)/3]. */
/* No reference anchor (the reference has no probes).                       */
/* ------------------------------------------------------------------------ */
int rco_probe_spacing(int level) { return 2 << level; }
int rco_num_dirs(int level) { return 4 << (2 * level); }

void rco_probe_grid(int W, int H, int level, int *px, int *py)
{
    int s = rco_probe_spacing(level);
    *px = (W + s - 1) / s;
    *py = (H + s - 1) / s;
}

int64_t rco_interval_t0(int level) { return (2 * ((1ll << (2 * level)) - 1)) / 3; }
int64_t rco_interval_t1(int level) { return (2 * ((1ll << (2 * level + 2)) - 1)) / 3; }

int64_t rco_level_rays(int W, int H, int level)
{
    int px, py;
    rco_probe_grid(W, H, level, &px, &py);
    return (int64_t)px * py * rco_num_dirs(level);
}

/* Smallest L whose top interval end 2(4^L - 1)/3 reaches the image diagonal. */
int rco_levels_for_coverage(int W, int H)
{
    int64_t diag2 = (int64_t)W * W + (int64_t)H * H;
    for (int L = 1; L <= RCO_MAX_LEVELS; ++L) {
        int64_t t = rco_interval_t1(L - 1);
        if (t * t >= diag2) return L;
    }
    return RCO_MAX_LEVELS;
}

int64_t rco_ray_index(int W, int H, int level, int px, int py, int k)
{
    int PX, PY;
    rco_probe_grid(W, H, level, &PX, &PY);
    (void)PY;
    return ((int64_t)py * PX + px) * rco_num_dirs(level) + k;
}

/* theta_k = 2 pi (k + 1/2) / N evaluated in double, rounded to fp32. */
void rco_dir_table_f32(int level, float *dirs)
{
    int N = rco_num_dirs(level);
    for (int k = 0; k < N; ++k) {
        double th = 2.0 * M_PI * ((double)k + 0.5) / (double)N;
        dirs[2 * k] = (float)cos(th);
        dirs[2 * k + 1] = (float)sin(th);
    }
}

/* ------------------------------------------------------------------------ */
/* Occupancy and jump flood.                                                 */
/* ------------------------------------------------------------------------ */
int rco_jfa_num_passes(int W, int H)
{
    int m = W > H ? W : H, P = 0;
    while ((1 << P) < m) ++P;
    return P;
}

/*
 * seed(p) = (y << 16) | x when p is occupied, else RCO_SEED_NONE.  Occupied:
 * any emissive channel > 0, or some in-bounds 4-neighbour q has
 * D(q) - D(p) > 0.05 D(p) evaluated in fp32 (p is the near side of a depth
 * silhouette).  Depth is the linear depth of the raster above; the reference's
 * depth target is Depth32Float (src/texture.rs:149).
 */
void rco_seed_init(int W, int H, const float *depth, const float *emissive, uint32_t *seed)
{
    const size_t plane = (size_t)W * (size_t)H;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            int occ = emissive[p] > 0.0f || emissive[plane + p] > 0.0f || emissive[2 * plane + p] > 0.0f;
            float dp = depth[p];
            float thr = 0.05f * dp;
            static const int nx[4] = {-1, 1, 0, 0}, ny[4] = {0, 0, -1, 1};
            for (int j = 0; j < 4 && !occ; ++j) {
                int qx = x + nx[j], qy = y + ny[j];
                if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                float diff = depth[(size_t)qy * W + qx] - dp;
                if (diff > thr) occ = 1;
            }
            seed[p] = occ ? (((uint32_t)y << 16) | (uint32_t)x) : RCO_SEED_NONE;
        }
    }
}

/* Order key: (d^2, seed_y, seed_x) lexicographic; the packed seed already
 * orders (y, x).  RCO_SEED_NONE maps to the largest key. */
static inline uint64_t jfa_key(uint32_t s, int x, int y)
{
    if (s == RCO_SEED_NONE) return ~0ull;
    int dx = (int)(s & 0xFFFFu) - x, dy = (int)(s >> 16) - y;
    uint32_t d2 = (uint32_t)(dx * dx + dy * dy);
    return ((uint64_t)d2 << 32) | s;
}

/* One jump-flood pass (Rong & Tan 2006): own seed and the 8 neighbours at
 * +-step, out-of-bounds neighbours skipped, minimum key wins. */
void rco_jfa_pass(int W, int H, int step, const uint32_t *in, uint32_t *out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            uint64_t best = jfa_key(in[(size_t)y * W + x], x, y);
            for (int dy = -1; dy <= 1; ++dy) {
                int qy = y + dy * step;
                if (qy < 0 || qy >= H) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    int qx = x + dx * step;
                    if ((dx == 0 && dy == 0) || qx < 0 || qx >= W) continue;
                    uint64_t k = jfa_key(in[(size_t)qy * W + qx], x, y);
                    if (k < best) best = k;
                }
            }
            out[(size_t)y * W + x] = best == ~0ull ? RCO_SEED_NONE : (uint32_t)best;
        }
    }
}

/* Steps 2^(P-1), ..., 2, 1 with P = ceil(log2 max(W,H)); result in seed. */
void rco_jfa(int W, int H, uint32_t *seed, uint32_t *tmp)
{
    int P = rco_jfa_num_passes(W, H);
    uint32_t *a = seed, *b = tmp;
    for (int p = P - 1; p >= 0; --p) {
        rco_jfa_pass(W, H, 1 << p, a, b);
        uint32_t *t = a; a = b; b = t;
    }
    if (a != seed) memcpy(seed, a, (size_t)W * H * sizeof(uint32_t));
}

static inline uint32_t seed_d2(uint32_t s, int x, int y)
{
    int dx = (int)(s & 0xFFFFu) - x, dy = (int)(s >> 16) - y;
    return (uint32_t)(dx * dx + dy * dy);
}

/* Distance the march reads: a conservative one-byte code of d = sqrt(d^2)
 * (rc_common.hpp of the product states the same rule): floor(d) below 128 px,
 * 128 + 16 floor((d - 128) / 16) below 2160 px, 2160 beyond and with no seed at
 * all.  The value never exceeds d, and is 0 exactly when d is. */
static uint32_t isqrt_u64(uint64_t n)
{
    uint64_t r = (uint64_t)sqrt((double)n);
    while (r * r > n) --r;
    while ((r + 1) * (r + 1) <= n) ++r;
    return (uint32_t)r;
}

double rco_sdf_value_d2(uint64_t d2)
{
    const uint32_t d = isqrt_u64(d2);
    if (d < 128) return (double)d;
    if (d < 2160) return 128.0 + 16.0 * (double)((d - 128) / 16);
    return 2160.0;
}

static double sdf_value_at(uint32_t s, int x, int y)
{
    return s == RCO_SEED_NONE ? 2160.0 : rco_sdf_value_d2(seed_d2(s, x, y));
}

void rco_sdf_f32(int W, int H, const uint32_t *seed, float *sdf)
{
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) sdf[(size_t)y * W + x] = (float)sdf_value_at(seed[(size_t)y * W + x], x, y);
}

void rco_sdf_f64(int W, int H, const uint32_t *seed, double *sdf)
{
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) sdf[(size_t)y * W + x] = sdf_value_at(seed[(size_t)y * W + x], x, y);
}

/*
 * Exact squared Euclidean distance transform of the occupancy (seed_init !=
 * none): d2[p] = min over occupied q of |p - q|^2, -1 when no pixel is
 * occupied.  Felzenszwalb & Huttenlocher, "Distance Transforms of Sampled
 * Functions" (Theory of Computing 8, 2012): per column the vertical distance
 * g to the nearest occupied pixel of that column, then per row the lower
 * envelope of the parabolas (x - q)^2 + g(q)^2 over the columns q with a
 * finite g, built left to right with a stack.  The envelope's breakpoints are
 * compared as exact integer cross products (no floating point), so the
 * result is exact for every image up to 16384 px a side.
 */
void rco_edt_d2(int W, int H, const uint32_t *seed_init, int64_t *d2)
{
    const int64_t NONE = -1;
    int64_t *g = (int64_t *)malloc((size_t)W * H * sizeof(int64_t));
#pragma omp parallel for schedule(static)
    for (int x = 0; x < W; ++x) {
        int64_t last = NONE;  /* row of the nearest occupied pixel above (or at) y */
        for (int y = 0; y < H; ++y) {
            if (seed_init[(size_t)y * W + x] != RCO_SEED_NONE) last = y;
            g[(size_t)y * W + x] = last == NONE ? NONE : y - last;
        }
        last = NONE;
        for (int y = H - 1; y >= 0; --y) {
            if (seed_init[(size_t)y * W + x] != RCO_SEED_NONE) last = y;
            if (last != NONE) {
                int64_t *v = &g[(size_t)y * W + x];
                if (*v == NONE || last - y < *v) *v = last - y;
            }
        }
    }
#pragma omp parallel
    {
        int *site = (int *)malloc((size_t)W * sizeof(int));
        int64_t *F = (int64_t *)malloc((size_t)W * sizeof(int64_t));  /* g(q)^2 + q^2 of stacked site */
#pragma omp for schedule(static)
        for (int y = 0; y < H; ++y) {
            const int64_t *gr = g + (size_t)y * W;
            int k = -1;
            for (int q = 0; q < W; ++q) {
                if (gr[q] == NONE) continue;
                const int64_t Fq = gr[q] * gr[q] + (int64_t)q * q;
                /* pop the top while the new parabola's crossing with it lies at
                 * or before the top's own crossing with the one below it:
                 * (Fq - Fv)/(q - v) <= (Fv - Fu)/(v - u), denominators > 0 */
                while (k >= 1) {
                    const int64_t v = site[k], u = site[k - 1];
                    if ((Fq - F[k]) * (v - u) <= (F[k] - F[k - 1]) * (q - v)) --k;
                    else break;
                }
                ++k;
                site[k] = q;
                F[k] = Fq;
            }
            int64_t *out = d2 + (size_t)y * W;
            if (k < 0) {
                for (int x = 0; x < W; ++x) out[x] = NONE;
                continue;
            }
            int j = 0;
            for (int x = 0; x < W; ++x) {
                /* advance while the next site's parabola is strictly lower at
                 * x: crossing (F[j+1] - F[j]) / (2 (s[j+1] - s[j])) < x */
                while (j < k && F[j + 1] - F[j] < 2 * (int64_t)x * (site[j + 1] - site[j])) ++j;
                const int64_t dx = x - site[j];
                out[x] = dx * dx + F[j] - (int64_t)site[j] * site[j];
            }
        }
        free(F);
        free(site);
    }
    free(g);
}

void rco_sdf_from_d2_f32(int W, int H, const int64_t *d2, float *sdf)
{
    for (size_t p = 0; p < (size_t)W * H; ++p) sdf[p] = d2[p] < 0 ? 2160.0f : (float)rco_sdf_value_d2((uint64_t)d2[p]);
}

void rco_sdf_from_d2_f64(int W, int H, const int64_t *d2, double *sdf)
{
    for (size_t p = 0; p < (size_t)W * H; ++p) sdf[p] = d2[p] < 0 ? 2160.0 : rco_sdf_value_d2((uint64_t)d2[p]);
}

/* Exact nearest seed under the same tie-break (checker for the jump flood). */
void rco_nearest_seed_bruteforce(int W, int H, const uint32_t *seed_init, uint32_t *nearest)
{
    size_t n = (size_t)W * H, ns = 0;
    uint32_t *list = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    for (size_t p = 0; p < n; ++p)
        if (seed_init[p] != RCO_SEED_NONE) list[ns++] = seed_init[p];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint64_t best = ~0ull;
            for (size_t j = 0; j < ns; ++j) {
                uint64_t k = jfa_key(list[j], x, y);
                if (k < best) best = k;
            }
            nearest[(size_t)y * W + x] = best == ~0ull ? RCO_SEED_NONE : (uint32_t)best;
        }
    free(list);
}

/*
 * Ground truth of the 2D transport the cascades approximate (a checker, not
 * a flavour of the method): from every listed pixel centre, n_dirs rays at
 * theta_k = 2 pi (k + 1/2) / n_dirs, each walked through every pixel it
 * crosses (Amanatides & Woo's grid traversal, FP64) until the first occupied
 * pixel (seed_init != none), which contributes its emission, or until it
 * leaves the image or passes distance t_max (no contribution).  A ray that
 * starts in an occupied pixel hits it.  out[3 * i + c] = mean over the
 * directions, channel c, for pixel pixels[i] (y * W + x; every pixel in order
 * when pixels is NULL, n_pix = W * H).
 */
void rco_bruteforce_irradiance(int W, int H, const uint32_t *seed_init, const float *emissive, int n_dirs,
                               double t_max, const int32_t *pixels, int n_pix, double *out)
{
    const size_t plane = (size_t)W * H;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n_pix; ++i) {
        const int p = pixels ? pixels[i] : i;
        const int x0 = p % W, y0 = p / W;
        double acc[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < n_dirs; ++k) {
            const double th = 2.0 * M_PI * ((double)k + 0.5) / (double)n_dirs;
            const double dx = cos(th), dy = sin(th);
            int ix = x0, iy = y0;
            const int sx = dx > 0 ? 1 : -1, sy = dy > 0 ? 1 : -1;
            const double adx = fabs(dx), ady = fabs(dy);
            /* the centre is half a pixel from either boundary */
            double tx = adx > 0 ? 0.5 / adx : INFINITY, ty = ady > 0 ? 0.5 / ady : INFINITY;
            const double ddx = adx > 0 ? 1.0 / adx : INFINITY, ddy = ady > 0 ? 1.0 / ady : INFINITY;
            for (;;) {
                const size_t q = (size_t)iy * W + ix;
                if (seed_init[q] != RCO_SEED_NONE) {
                    acc[0] += emissive[q];
                    acc[1] += emissive[plane + q];
                    acc[2] += emissive[2 * plane + q];
                    break;
                }
                if (tx <= ty) {
                    if (tx > t_max) break;
                    ix += sx;
                    tx += ddx;
                } else {
                    if (ty > t_max) break;
                    iy += sy;
                    ty += ddy;
                }
                if (ix < 0 || iy < 0 || ix >= W || iy >= H) break;
            }
        }
        for (int c = 0; c < 3; ++c) out[3 * (size_t)i + c] = acc[c] / (double)n_dirs;
    }
}

void rco_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int rco_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- the two arithmetic flavours of march / merge / gather / run ---------- */
#define RCO_REAL double
#define RCO_FN(name) name##_f64
#define RCO_IS_F32 0
#include "rco_cascade.inc"
#undef RCO_REAL
#undef RCO_FN
#undef RCO_IS_F32

#define RCO_REAL float
#define RCO_FN(name) name##_f32
#define RCO_IS_F32 1
#include "rco_cascade.inc"
#undef RCO_REAL
#undef RCO_FN
#undef RCO_IS_F32
