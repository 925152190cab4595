// rc_norms.hip -- error norms of a fp32 output against a reference, reduced
// in FP64 on the GPU (the only FP64 arithmetic of the GI path).
//
// out[0] = sum (x - r)^2, out[1] = sum r^2, out[2] = max |x - r|, with
// rel-L2 = sqrt(out[0] / out[1]).  Two deterministic stages: every block
// reduces a grid-stride share into a partial (wave DPP-free shuffles, then
// LDS), and one block reduces the partials in a fixed order, so repeated
// calls return the same bits.
#include "rc_internal.hpp"

namespace rc {

constexpr int kNormBlocks = 1024;

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <bool REF64>
__global__ __launch_bounds__(kBlock) void k_norm_partial(const float *__restrict__ x, const void *__restrict__ ref,
                                                         size_t n, double *__restrict__ partial)
{
    double s_d = 0.0, s_r = 0.0, m = 0.0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const double r = REF64 ? ((const double *)ref)[i] : (double)((const float *)ref)[i];
        const double d = (double)x[i] - r;
        s_d = fma(d, d, s_d);
        s_r = fma(r, r, s_r);
        m = fmax(m, fabs(d));
    }
    s_d = wave_sum(s_d);
    s_r = wave_sum(s_r);
    m = wave_max(m);
    __shared__ double sh[3][kBlock / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = s_d;
        sh[1][w] = s_r;
        sh[2][w] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int j = 0; j < kBlock / 64; ++j) {
            a += sh[0][j];
            b += sh[1][j];
            c = fmax(c, sh[2][j]);
        }
        partial[3 * blockIdx.x + 0] = a;
        partial[3 * blockIdx.x + 1] = b;
        partial[3 * blockIdx.x + 2] = c;
    }
}

__global__ __launch_bounds__(kBlock) void k_norm_final(const double *__restrict__ partial, int blocks,
                                                       double *__restrict__ out)
{
    double s_d = 0.0, s_r = 0.0, m = 0.0;
    for (int i = threadIdx.x; i < blocks; i += kBlock) {
        s_d += partial[3 * i + 0];
        s_r += partial[3 * i + 1];
        m = fmax(m, partial[3 * i + 2]);
    }
    s_d = wave_sum(s_d);
    s_r = wave_sum(s_r);
    m = wave_max(m);
    __shared__ double sh[3][kBlock / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = s_d;
        sh[1][w] = s_r;
        sh[2][w] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int j = 0; j < kBlock / 64; ++j) {
            a += sh[0][j];
            b += sh[1][j];
            c = fmax(c, sh[2][j]);
        }
        out[0] = a;
        out[1] = b;
        out[2] = c;
    }
}

size_t norm_workspace_bytes() { return (size_t)kNormBlocks * 3 * sizeof(double); }

hipError_t launch_error_norms(const float *x, const void *ref, bool ref_f64, size_t n, double *workspace,
                              double *out, hipStream_t s)
{
    size_t want = (n + kBlock - 1) / kBlock;
    const int blocks = (int)(want < (size_t)kNormBlocks ? (want ? want : 1) : kNormBlocks);
    if (ref_f64) hipLaunchKernelGGL(k_norm_partial<true>, dim3(blocks), dim3(kBlock), 0, s, x, ref, n, workspace);
    else hipLaunchKernelGGL(k_norm_partial<false>, dim3(blocks), dim3(kBlock), 0, s, x, ref, n, workspace);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(kBlock), 0, s, workspace, blocks, out);
    return hipGetLastError();
}

}  // namespace rc
