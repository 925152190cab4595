// rc_mlp.hip -- the surrogate MLP: BF16 MFMA forward (inference and
// training), backward, gradient reduction and Adam.
//
// Network (the project's own; the reference has no learned part):
// 64 -> 64 -> 64 -> 64 -> 64, ReLU on the hidden layers.  Outputs 0..47 are
// the level-1 far-field radiance (16 directions x rgb) in normalised units,
// output 48 the log of the predicted per-sample squared error (the per-tile
// fallback signal), outputs 49..63 padding.  Layer l computes
// z = bf16(a) . bf16(W) + b with FP32 accumulation on
// v_mfma_f32_32x32x16_bf16; the backward rounds each upstream gradient to
// bf16 before its two products (dW and dA) and keeps bias gradients in FP32.
//
// MFMA operand layout (32x32x16, wave64): lane l supplies row (l % 32) of A
// and column (l % 32) of B, each as 8 consecutive K elements starting at
// 8 * (l / 32); accumulator register v of lane l holds
// C[(v / 4) * 8 + (l / 32) * 4 + v % 4][l % 32].  Every operand is staged so
// that its 8 K elements are 16 contiguous bytes in LDS or global memory.
#include <string.h>

#include <vector>

#include "rc_internal.hpp"

namespace rc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---- parameter layout ---------------------------------------------------------
// FP32 master parameters: for l = 0..3, W_l [K_l][64] (input-major) then
// b_l [64].  bf16 copies: Wt_l [64][K_l] (output-major, forward B operand) and
// Wk_l [K_l][64] (input-major, backward B operand).
__host__ __device__ constexpr int mlp_k(int l) { return l == 0 ? kMlpIn : kMlpHid; }
__host__ __device__ constexpr int mlp_w_off(int l) { return l == 0 ? 0 : kMlpIn * 64 + (l - 1) * 64 * 64; }
__host__ __device__ constexpr int mlp_p_w(int l) { return mlp_w_off(l) + l * 64; }  // in the fp32 master vector
__host__ __device__ constexpr int mlp_p_b(int l) { return mlp_p_w(l) + mlp_k(l) * 64; }
static_assert(mlp_p_b(3) + 64 == kMlpParams, "parameter count");

__device__ inline uint16_t bf16_bits(float f) { return (uint16_t)f32_to_bf16_rne(f); }

// LDS row strides (elements) with 8 bf16 of padding: 16-byte aligned rows that
// do not all start on the same bank.
constexpr int kStrideIn = kMlpIn + 8;    // 56
constexpr int kStrideHid = kMlpHid + 8;  // 72
constexpr int kStrideT = 32 + 8;         // transposed activations [feature][sample]

__device__ inline bf16x8 ld8(const uint16_t *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// acc[nt] = A (32 x K, row stride sa) . B, B column n = nt*32 + (l%32) read as
// 8 consecutive K elements at b + n*sb + k.
template <int K>
__device__ inline void mfma_rows32(const uint16_t *a, int sa, const uint16_t *b, int sb, floatx16 acc[2])
{
    const int lane = threadIdx.x & 63, r = lane & 31, kh = (lane >> 5) * 8;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        floatx16 c = {};
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 16)
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(a + r * sa + k0 + kh), ld8(b + (nt * 32 + r) * sb + k0 + kh), c,
                                                        0, 0, 0);
        acc[nt] = c;
    }
}

__device__ inline int acc_row(int v, int lane) { return ((v >> 2) << 3) + ((lane >> 5) << 2) + (v & 3); }

// Forward of one wave's 32 rows.  x: normalised features as bf16 [32][56] in
// LDS; h: three hidden buffers [32][72]; wt: bf16 Wt_l (LDS or global);
// bias: FP32.  Returns the output layer's accumulators (+bias) in z[2].
// hT (optional): transposed copies [64][40] of the inputs of layers 2..4.
__device__ inline void mlp_forward32(const uint16_t *x, uint16_t *h1, uint16_t *h2, uint16_t *h3,
                                     const uint16_t *const wt[4], const float *bias, floatx16 z[2],
                                     uint16_t *hT1 = nullptr, uint16_t *hT2 = nullptr, uint16_t *hT3 = nullptr)
{
    const int lane = threadIdx.x & 63;
    floatx16 acc[2];
    const uint16_t *in = x;
    int sin = kStrideIn;
    uint16_t *outs[3] = {h1, h2, h3};
    uint16_t *outsT[3] = {hT1, hT2, hT3};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        if (l == 0) mfma_rows32<kMlpIn>(in, sin, wt[0], kMlpIn, acc);
        else mfma_rows32<kMlpHid>(in, sin, wt[l], kMlpHid, acc);
        if (l == 3) {
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const float bb = bias[3 * 64 + nt * 32 + (lane & 31)];
#pragma unroll
                for (int v = 0; v < 16; ++v) z[nt][v] = acc[nt][v] + bb;
            }
            break;
        }
        __syncthreads();  // previous readers of the output buffer are done (one wave: a no-op barrier)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int col = nt * 32 + (lane & 31);
            const float bb = bias[l * 64 + col];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = acc_row(v, lane);
                const uint16_t hv = bf16_bits(fmaxf(acc[nt][v] + bb, 0.0f));
                outs[l][row * kStrideHid + col] = hv;
                if (outsT[l]) outsT[l][col * kStrideT + row] = hv;
            }
        }
        __syncthreads();
        in = outs[l];
        sin = kStrideHid;
    }
}

// ---- inference over a feature matrix -------------------------------------------
// features [n][64] (already transformed, standardised here) -> z [n][64]
// fp32.  Each wave takes a 32-row block per iteration; the iteration count is
// the same for every wave of the workgroup (a wave past the end computes on
// zero rows and stores nothing), so the barriers are uniform.
__global__ __launch_bounds__(kBlock) void k_mlp_infer(MlpDev m, const float *__restrict__ feats, int n,
                                                      float *__restrict__ out)
{
    __shared__ __attribute__((aligned(16))) uint16_t wt_s[kMlpWt];
    __shared__ float bias_s[4 * 64], mean_s[kMlpIn], istd_s[kMlpIn];
    __shared__ __attribute__((aligned(16))) uint16_t buf[4][32 * kStrideIn + 3 * 32 * kStrideHid];
    for (int i = threadIdx.x; i < kMlpWt; i += kBlock) wt_s[i] = m.wt[i];
    for (int i = threadIdx.x; i < 4 * 64; i += kBlock) bias_s[i] = m.bias[i];
    for (int i = threadIdx.x; i < kMlpIn; i += kBlock) {
        mean_s[i] = m.in_mean[i];
        istd_s[i] = m.in_istd[i];
    }
    __syncthreads();
    const uint16_t *wt[4] = {wt_s, wt_s + 64 * kMlpIn, wt_s + 64 * kMlpIn + 64 * kMlpHid,
                             wt_s + 64 * kMlpIn + 2 * 64 * kMlpHid};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t *x = buf[wave], *h1 = x + 32 * kStrideIn, *h2 = h1 + 32 * kStrideHid, *h3 = h2 + 32 * kStrideHid;
    const int blocks = (n + 31) / 32;
    for (int base = blockIdx.x * 4; base < blocks; base += gridDim.x * 4) {
        const int row0 = (base + wave) * 32;
        for (int i = lane; i < 32 * kMlpIn; i += 64) {
            const int r = i / kMlpIn, c = i % kMlpIn;
            const float v = row0 + r < n ? (feats[(size_t)(row0 + r) * kMlpIn + c] - mean_s[c]) * istd_s[c] : 0.0f;
            x[r * kStrideIn + c] = bf16_bits(v);
        }
        __syncthreads();
        floatx16 z[2];
        mlp_forward32(x, h1, h2, h3, wt, bias_s, z);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = row0 + acc_row(v, lane);
                if (row < n) out[(size_t)row * kMlpOut + nt * 32 + (lane & 31)] = z[nt][v];
            }
        __syncthreads();
    }
}

// ---- training: forward + backward, per-workgroup gradient partials -----------
// One wave per workgroup; the workgroup takes `nb` 32-sample blocks of the
// batch (sample ids from idx) and writes the FP32 sum of their gradients to
// partial[blockIdx.x][kMlpParams] (bias gradients included) and its loss sum.
// Targets are normalised [n][48].  Weights are read from global memory (L2).
__global__ __launch_bounds__(64) void k_mlp_train(MlpDev m, const float *__restrict__ feats,
                                                  const float *__restrict__ targets, const int *__restrict__ idx,
                                                  int batch, int nb, float *__restrict__ partial,
                                                  float *__restrict__ loss_part)
{
    __shared__ __attribute__((aligned(16))) uint16_t x[32 * kStrideIn], h1[32 * kStrideHid], h2[32 * kStrideHid],
        h3[32 * kStrideHid];
    __shared__ __attribute__((aligned(16))) uint16_t xT[64 * kStrideT], hT1[64 * kStrideT], hT2[64 * kStrideT],
        hT3[64 * kStrideT];
    __shared__ __attribute__((aligned(16))) uint16_t g[32 * kStrideHid], gT[64 * kStrideT];
    __shared__ float zs[32][kMlpOut + 1];
    __shared__ float tg[32][kMlpRad + 1];
    __shared__ float gsum[64];
    const int lane = threadIdx.x;
    const uint16_t *wt[4] = {m.wt, m.wt + 64 * kMlpIn, m.wt + 64 * kMlpIn + 64 * kMlpHid,
                             m.wt + 64 * kMlpIn + 2 * 64 * kMlpHid};
    const uint16_t *wk[4] = {m.wk, m.wk + 64 * kMlpIn, m.wk + 64 * kMlpIn + 64 * kMlpHid,
                             m.wk + 64 * kMlpIn + 2 * 64 * kMlpHid};
    float *part = partial + (size_t)blockIdx.x * kMlpParams;
    float loss_acc = 0.0f;
    const float inv_bn = 1.0f / ((float)batch * kMlpRad), inv_b = 1.0f / (float)batch;

    for (int bi = 0; bi < nb; ++bi) {
        const int s0 = (blockIdx.x * nb + bi) * 32;
        // features (normalised) -> x and xT; targets -> tg.  Rows past the
        // batch are zero features with zero gradient.
        for (int i = lane; i < 32 * kMlpIn; i += 64) {
            const int r = i / kMlpIn, c = i % kMlpIn;
            const bool ok = s0 + r < batch;
            const int sid = ok ? idx[s0 + r] : 0;
            const float v = ok ? (feats[(size_t)sid * kMlpIn + c] - m.in_mean[c]) * m.in_istd[c] : 0.0f;
            const uint16_t b = bf16_bits(v);
            x[r * kStrideIn + c] = b;
            xT[c * kStrideT + r] = b;
            tg[r][c] = ok ? targets[(size_t)sid * kMlpRad + c] : 0.0f;
        }
        for (int i = lane; i < (64 - kMlpIn) * 32; i += 64) xT[(kMlpIn + i / 32) * kStrideT + (i % 32)] = 0;  // pad rows kMlpIn..63
        __syncthreads();
        floatx16 z[2];
        mlp_forward32(x, h1, h2, h3, wt, m.bias, z, hT1, hT2, hT3);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int v = 0; v < 16; ++v) zs[acc_row(v, lane)][nt * 32 + (lane & 31)] = z[nt][v];
        __syncthreads();

        // dL/dz4 per row: lanes 0..31 own one sample each (FP32, deterministic)
        if (lane < 32) {
            const bool ok = s0 + lane < batch;
            float se = 0.0f;
            for (int c = 0; c < kMlpRad; ++c) {
                const float d = zs[lane][c] - tg[lane][c];
                se = fmaf(d, d, se);
            }
            const float e_t = logf(se * (1.0f / kMlpRad) + kMlpErrEps);
            const float de = zs[lane][kMlpErr] - e_t;
            if (ok) loss_acc += se * inv_bn + kMlpErrWeight * de * de * inv_b;
            for (int c = 0; c < kMlpOut; ++c) {
                float gv = 0.0f;
                if (ok && c < kMlpRad) gv = 2.0f * (zs[lane][c] - tg[lane][c]) * inv_bn;
                if (ok && c == kMlpErr) gv = kMlpErrWeight * 2.0f * de * inv_b;
                zs[lane][c] = gv;  // reuse as the FP32 gradient
            }
        }
        __syncthreads();

        // backward, layer 3 .. 0
        for (int l = 3; l >= 0; --l) {
            const int K = mlp_k(l);
            const uint16_t *aTl = l == 0 ? xT : (l == 1 ? hT1 : (l == 2 ? hT2 : hT3));
            const uint16_t *arl = l == 1 ? h1 : (l == 2 ? h2 : h3);
            // G = bf16(g) row-major [32][64] and transposed [64][40]; bias grad = column sums
            for (int i = lane; i < 32 * 64; i += 64) {
                const int r = i >> 6, c = i & 63;
                const uint16_t b = bf16_bits(zs[r][c]);
                g[r * kStrideHid + c] = b;
                gT[c * kStrideT + r] = b;
            }
            {
                float s = 0.0f;
                for (int r = 0; r < 32; ++r) s += zs[r][lane];
                gsum[lane] = s;
            }
            __syncthreads();
            part[mlp_p_b(l) + lane] = (bi == 0 ? 0.0f : part[mlp_p_b(l) + lane]) + gsum[lane];
            // dW^T [64 out][K in] = G^T (64 x 32 samples) . A_l (32 samples x K):
            // A operand rows = outputs from gT, B operand columns = inputs from aT
            for (int mt = 0; mt < 2; ++mt) {
                floatx16 acc[2];
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) {
                    floatx16 c = {};
                    const int r = lane & 31, kh = (lane >> 5) * 8;
#pragma unroll
                    for (int k0 = 0; k0 < 32; k0 += 16)
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            ld8(gT + (mt * 32 + r) * kStrideT + k0 + kh), ld8(aTl + (nt * 32 + r) * kStrideT + k0 + kh),
                            c, 0, 0, 0);
                    acc[nt] = c;
                }
                // element (out o, in i) of dW^T -> master layout W_l[i][o]
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int o = mt * 32 + acc_row(v, lane), in = nt * 32 + (lane & 31);
                        if (in < K) {
                            float *p = part + mlp_p_w(l) + in * 64 + o;
                            *p = (bi == 0 ? 0.0f : *p) + acc[nt][v];
                        }
                    }
            }
            if (l == 0) break;
            // dA_{l} = G (32 x 64) . W_l^T (64 x K): B column n = input, K dim = outputs,
            // read from Wk_l [K][64] (input-major), then the ReLU mask of the forward
            floatx16 da[2];
            {
                const int r = lane & 31, kh = (lane >> 5) * 8;
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) {
                    floatx16 c = {};
#pragma unroll
                    for (int k0 = 0; k0 < 64; k0 += 16)
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(g + r * kStrideHid + k0 + kh),
                                                                    ld8(wk[l] + (nt * 32 + r) * 64 + k0 + kh), c, 0, 0, 0);
                    da[nt] = c;
                }
            }
            __syncthreads();
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int row = acc_row(v, lane), col = nt * 32 + (lane & 31);
                    // the forward's ReLU output is > 0 exactly where z > 0
                    const uint16_t hv = arl[row * kStrideHid + col];
                    zs[row][col] = (hv != 0 && (hv & 0x8000u) == 0) ? da[nt][v] : 0.0f;
                }
            __syncthreads();
        }
        __syncthreads();
    }
    // per-workgroup loss (lanes 0..31 hold row losses)
    for (int o = 32; o >= 1; o >>= 1) loss_acc += __shfl_xor(loss_acc, o, 64);
    if (lane == 0) loss_part[blockIdx.x] = loss_acc;
}

// Sum the workgroup partials in a fixed order (deterministic) into grad.
__global__ __launch_bounds__(kBlock) void k_mlp_reduce(const float *__restrict__ partial, int parts,
                                                       float *__restrict__ grad)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= kMlpParams) return;
    float s = 0.0f;
    for (int p = 0; p < parts; ++p) s += partial[(size_t)p * kMlpParams + i];
    grad[i] = s;
}

// Adam (Kingma & Ba 2015, bias-corrected) on the FP32 master parameters, then
// refresh the bf16 copies Wt (output-major) and Wk (input-major).
__global__ __launch_bounds__(kBlock) void k_mlp_adam(MlpDev m, const float *__restrict__ grad, float lr, float b1,
                                                     float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= kMlpParams) return;
    const float gi = grad[i];
    const float mi = b1 * m.adam_m[i] + (1.0f - b1) * gi;
    const float vi = b2 * m.adam_v[i] + (1.0f - b2) * gi * gi;
    m.adam_m[i] = mi;
    m.adam_v[i] = vi;
    const float p = m.master[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    m.master[i] = p;
    // scatter to the bf16 copies
    for (int l = 0; l < 4; ++l) {
        const int K = mlp_k(l), w0 = mlp_p_w(l), b0 = mlp_p_b(l);
        if (i >= w0 && i < b0) {
            const int in = (i - w0) / 64, o = (i - w0) % 64;
            const uint16_t b = bf16_bits(p);
            m.wt[mlp_w_off(l) + o * K + in] = b;
            m.wk[mlp_w_off(l) + in * 64 + o] = b;
        } else if (i >= b0 && i < b0 + 64) {
            m.bias[l * 64 + (i - b0)] = p;
        }
    }
}

// Refresh the bf16 copies and biases from the master vector (after a load).
__global__ __launch_bounds__(kBlock) void k_mlp_refresh(MlpDev m)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= kMlpParams) return;
    const float p = m.master[i];
    for (int l = 0; l < 4; ++l) {
        const int K = mlp_k(l), w0 = mlp_p_w(l), b0 = mlp_p_b(l);
        if (i >= w0 && i < b0) {
            const int in = (i - w0) / 64, o = (i - w0) % 64;
            m.wt[mlp_w_off(l) + o * K + in] = bf16_bits(p);
            m.wk[mlp_w_off(l) + in * 64 + o] = bf16_bits(p);
        } else if (i >= b0 && i < b0 + 64) {
            m.bias[l * 64 + (i - b0)] = p;
        }
    }
}

// Deterministic sum of n floats (one workgroup, fixed order per lane then a
// fixed tree).
__global__ __launch_bounds__(kBlock) void k_sum_f32(const float *__restrict__ x, int n, float *__restrict__ out)
{
    __shared__ float s[kBlock];
    float a = 0.0f;
    for (int i = threadIdx.x; i < n; i += kBlock) a += x[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = kBlock / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

hipError_t launch_sum_f32(const float *x, int n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_f32, dim3(1), dim3(kBlock), 0, s, x, n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const MlpDev &m, const float *feats, int n, float *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int blocks32 = (n + 31) / 32;
    int grid = (blocks32 + 3) / 4;
    if (grid > 2048) grid = 2048;
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(kBlock), 0, s, m, feats, n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_train(const MlpDev &m, const float *feats, const float *targets, const int *idx, int batch,
                            int nb, float *partial, float *loss_part, int parts, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_train, dim3(parts), dim3(64), 0, s, m, feats, targets, idx, batch, nb, partial, loss_part);
    return hipGetLastError();
}

hipError_t launch_mlp_reduce(const float *partial, int parts, float *grad, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + kBlock - 1) / kBlock), dim3(kBlock), 0, s, partial, parts, grad);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpDev &m, const float *grad, float lr, float b1, float b2, float eps, float c1,
                           float c2, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m, grad, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_refresh(const MlpDev &m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_refresh, dim3((kMlpParams + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m);
    return hipGetLastError();
}

}  // namespace rc
