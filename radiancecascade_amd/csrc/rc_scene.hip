// rc_scene.hip -- seeded synthetic box scenes and their analytic G-buffer.
//
// This is the producer of the benchmark input, not part of the timed pass.
// The reference loads OBJ scenes (src/primitives.rs:122-175) and has no
// emissive material (src/primitives.rs:75-83), so the scene is the project's
// own: a unit-cube room (albedo 0.75) with axis-aligned boxes, camera at the
// room centre looking along +z with up +y in the reference's look_to_rh frame
// (src/camera.rs:43-52) and a 60 degree vertical field of view as the fovy of
// perspective_rh (src/camera.rs:77-79).  Depth is the linear view-space depth.
//
// The raster is FP64 with contraction off (the library is compiled with
// -ffp-contract=off): each pixel's result is a fixed sequence of correctly
// rounded IEEE operations, so it is reproducible bit for bit on any IEEE host.
#include "rc_internal.hpp"

namespace rc {

static uint64_t splitmix64_next(uint64_t &state)
{
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double unit_draw(uint64_t &state) { return (double)(splitmix64_next(state) >> 11) * 0x1.0p-53; }

// Per box: centre x,y,z ~ U[0,1); edges ~ 0.1 + 0.3 U; albedo ~ 0.2 + 0.7 U.
// A box overlapping the open clear zone (0.35,0.65)^3 around the camera is
// redrawn.  Then the first n_emissive boxes draw radiance 1 + 9 U per channel.
void scene_generate(uint64_t seed, int n_boxes, int n_emissive, rc_box *boxes)
{
    uint64_t st = seed;
    for (int b = 0; b < n_boxes; ++b) {
        bool blocked = true;
        while (blocked) {
            double c[3], e[3], a[3];
            for (double &v : c) v = unit_draw(st);
            for (double &v : e) v = 0.1 + 0.3 * unit_draw(st);
            for (double &v : a) v = 0.2 + 0.7 * unit_draw(st);
            blocked = true;
            for (int j = 0; j < 3; ++j) {
                boxes[b].lo[j] = c[j] - 0.5 * e[j];
                boxes[b].hi[j] = c[j] + 0.5 * e[j];
                boxes[b].albedo[j] = a[j];
                boxes[b].emit[j] = 0.0;
                blocked = blocked && boxes[b].lo[j] < 0.65 && 0.35 < boxes[b].hi[j];
            }
        }
    }
    for (int b = 0; b < n_emissive; ++b)
        for (double &v : boxes[b].emit) v = 1.0 + 9.0 * unit_draw(st);
}

// One thread per pixel; the box list (12 doubles per box) is read by every
// thread, so it sits in LDS.
__global__ __launch_bounds__(kBlock) void k_raster(const double *__restrict__ boxes, int n, int W, int H,
                                                   float *__restrict__ depth, float *__restrict__ albedo,
                                                   float *__restrict__ emissive)
{
    __shared__ double sb[RC_MAX_BOXES * 12];
    for (int i = threadIdx.x; i < n * 12; i += blockDim.x) sb[i] = boxes[i];
    __syncthreads();

    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;

    const double th = 1.0 / sqrt(3.0);
    const double aspect = (double)W / (double)H;
    const double o = 0.5;
    const double ndc_x = (2.0 * ((double)x + 0.5)) / (double)W - 1.0;
    const double ndc_y = 1.0 - (2.0 * ((double)y + 0.5)) / (double)H;
    double d[3];
    d[0] = -((ndc_x * th) * aspect);
    d[1] = ndc_y * th;
    d[2] = 1.0;

    double best = __builtin_huge_val();
    for (int a = 0; a < 3; ++a) {
        double t;
        if (d[a] > 0.0) t = (1.0 - o) / d[a];
        else if (d[a] < 0.0) t = (0.0 - o) / d[a];
        else continue;
        if (t < best) best = t;
    }
    int obj = -1;
    for (int b = 0; b < n; ++b) {
        const double *bx = sb + 12 * b;
        double tn = -__builtin_huge_val(), tf = __builtin_huge_val();
        for (int a = 0; a < 3; ++a) {
            double t1 = (bx[a] - o) / d[a];
            double t2 = (bx[3 + a] - o) / d[a];
            if (t1 > t2) { double s = t1; t1 = t2; t2 = s; }
            if (t1 > tn) tn = t1;
            if (t2 < tf) tf = t2;
        }
        if (tn <= tf && tn > 0.0 && tn < best) { best = tn; obj = b; }
    }
    const size_t p = (size_t)y * W + x, plane = (size_t)W * H;
    depth[p] = (float)(best * d[2]);
    for (int c = 0; c < 3; ++c) {
        if (albedo) albedo[c * plane + p] = obj < 0 ? 0.75f : (float)sb[12 * obj + 6 + c];
        emissive[c * plane + p] = obj < 0 ? 0.0f : (float)sb[12 * obj + 9 + c];
    }
}

hipError_t launch_raster(const double *boxes_dev, int n, int W, int H, float *depth, float *albedo, float *emissive,
                         hipStream_t stream)
{
    dim3 grid((W + 63) / 64, (H + 3) / 4);
    hipLaunchKernelGGL(k_raster, grid, dim3(kBlock), 0, stream, boxes_dev, n, W, H, depth, albedo, emissive);
    return hipGetLastError();
}

}  // namespace rc
