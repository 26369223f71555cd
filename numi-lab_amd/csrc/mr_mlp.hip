// Warm-start MLP training kernels (gfx950): forward with recorded
// activations, MSE loss gradient, backward (matrix-core GEMMs over the batch)
// and Adam. All matrix products take BF16 operands and accumulate in FP32 on
// v_mfma_f32_16x16x32_bf16; every reduction runs in a fixed order, so a
// training run repeats bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mr_mlp_host.h"
#include "mr_mlp_tile.h"

namespace {

using mr::kMlpHidden;
using mr::kMlpIn;
using mr::kMlpOut;
using mr::kMlpRows;
using mr::mlp_bf16x8;
using mr::mlp_f32x4;
using mr::MlpView;

extern "C" __global__ void __launch_bounds__(64) mr_mlp_forward_kernel(MlpView m, const float* x, uint32_t n,
                                                                       float* y_hat, mr::MlpTrace trace,
                                                                       uint32_t has_trace) {
    __shared__ __align__(16) unsigned char lds[mr::kMlpLdsBytes];
    const uint32_t lane = threadIdx.x;
    const uint32_t row0 = blockIdx.x * kMlpRows;
    const uint32_t rows = n - row0 < kMlpRows ? n - row0 : kMlpRows;
    const mr::MlpLds s = mr::mlp_lds(lds);
    for (uint32_t idx = lane; idx < kMlpRows * kMlpIn; idx += 64u) {
        const uint32_t row = idx / kMlpIn, f = idx % kMlpIn;
        s.x[idx] = row < rows ? (x[(size_t)(row0 + row) * kMlpIn + f] - m.in_shift[f]) * m.in_scale[f] : 0.0f;
    }
    __syncthreads();
    float y[4];
    mr::mlp_tile_forward(m, s, lane, y, has_trace != 0u ? &trace : nullptr, row0, rows);
    const uint32_t col = lane & 15u;
    if (col < kMlpOut) {
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) {
            const uint32_t row = (lane >> 4) * 4u + r;
            if (row < rows) y_hat[(size_t)(row0 + row) * kMlpOut + col] = y[r];
        }
    }
}

// d = y_hat - t / out_scale. With delta = 0 the loss is mean(d^2) and
// dy = 2 d / (3 n); with delta > 0 it is the Huber loss, mean(h(d)) with
// h(d) = d^2 / 2 for |d| <= delta, delta (|d| - delta / 2) beyond, and
// dy = clamp(d, -delta, delta) / (3 n). Per-block partial sums of the loss
// terms are summed by mr_mlp_sum_kernel in block order.
extern "C" __global__ void __launch_bounds__(256) mr_mlp_mse_kernel(const float* y_hat, const float* t,
                                                                    const float* out_scale, uint32_t n, float delta,
                                                                    float* dy, float* partial) {
    __shared__ float red[256];
    const uint32_t total = n * kMlpOut;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    float se = 0.0f;
    if (i < total) {
        const float d = y_hat[i] - t[i] / out_scale[i % kMlpOut];
        if (delta > 0.0f) {
            const float ad = fabsf(d);
            const float g = ad <= delta ? d : (d > 0.0f ? delta : -delta);
            dy[i] = g / (float)total;
            se = ad <= delta ? 0.5f * d * d : delta * (ad - 0.5f * delta);
        } else {
            dy[i] = 2.0f * d / (float)total;
            se = d * d;
        }
    }
    red[threadIdx.x] = se;
    __syncthreads();
    for (uint32_t w = 128u; w > 0u; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0u) partial[blockIdx.x] = red[0];
}

extern "C" __global__ void __launch_bounds__(256) mr_mlp_sum_kernel(const float* partial, uint32_t count,
                                                                    float scale, float* out) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (uint32_t i = threadIdx.x; i < count; i += 256u) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t w = 128u; w > 0u; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0u) *out = red[0] * scale;
}

// dz2[m][j] = (sum_o dy[m][o] W3[o][j]) * gelu'(z2[m][j])
extern "C" __global__ void __launch_bounds__(256) mr_mlp_out_backward_kernel(const float* dy, const __bf16* w3,
                                                                             const float* z2, uint32_t n,
                                                                             float* dz2) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= (size_t)n * kMlpHidden) return;
    const size_t m = i / kMlpHidden;
    const uint32_t j = (uint32_t)(i % kMlpHidden);
    float g = 0.0f;
#pragma unroll
    for (uint32_t o = 0; o < kMlpOut; ++o) g += dy[m * kMlpOut + o] * (float)w3[(size_t)o * kMlpHidden + j];
    dz2[i] = g * mr::mlp_gelu_grad(z2[i]);
}

template <class T> __device__ __forceinline__ __bf16 as_bf16(T v) { return (__bf16)v; }

// C[i][j] = sum_p A(i, p) B(p, j) with strided operands (FP32 or BF16
// sources, fed to the matrix cores as BF16) and FP32 accumulation in p order
// 32 at a time. One wavefront per 16x16 tile; grid (ceil(N/16), ceil(M/16)).
// With gate != null the result is multiplied by gelu'(gate[i][j]).
template <class TA, class TB>
__global__ void __launch_bounds__(64) mlp_gemm_kernel(uint32_t M, uint32_t N, uint32_t K, const TA* A, uint32_t a_si,
                                                      uint32_t a_sp, const TB* B, uint32_t b_sp, uint32_t b_sj,
                                                      float* C, uint32_t c_ld, const float* gate) {
    const uint32_t lane = threadIdx.x;
    const uint32_t ar = lane & 15u, kg = lane >> 4;
    const uint32_t i0 = blockIdx.y * 16u, j0 = blockIdx.x * 16u;
    const uint32_t ai = i0 + ar, bj = j0 + ar;
    mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t p0 = 0; p0 < K; p0 += 32u) {
        mlp_bf16x8 a, b;
#pragma unroll
        for (uint32_t e = 0; e < 8u; ++e) {
            const uint32_t p = p0 + kg * 8u + e;
            a[e] = (ai < M && p < K) ? as_bf16(A[(size_t)ai * a_si + (size_t)p * a_sp]) : (__bf16)0.0f;
            b[e] = (bj < N && p < K) ? as_bf16(B[(size_t)p * b_sp + (size_t)bj * b_sj]) : (__bf16)0.0f;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (uint32_t r = 0; r < 4u; ++r) {
        const uint32_t i = i0 + kg * 4u + r, j = j0 + ar;
        if (i < M && j < N) {
            float v = acc[r];
            if (gate != nullptr) v *= mr::mlp_gelu_grad(gate[(size_t)i * c_ld + j]);
            C[(size_t)i * c_ld + j] = v;
        }
    }
}

// Weight gradient C[i][j] = sum_p A[p][i] B[p][j] over a long row axis p
// (the batch): A FP32 [K][a_ld], B [K][b_ld] (BF16 or FP32), both fed to the
// matrix cores as BF16 with FP32 accumulation. One workgroup of 16
// wavefronts per 16x16 output tile: wavefront w takes the 32-row chunks
// w, w + 16, w + 32, ... of p, and the 16 partial tiles are summed through
// LDS in wavefront order, so the result does not depend on scheduling. With
// bias != null, the workgroups of the first column tile also return
// bias[i] = sum_p A[p][i] from the FP32 values (the bias gradient), summed in
// the same fixed order. Grid (ceil(N/16), ceil(M/16)), block 1024.
constexpr uint32_t kWgradWaves = 16u;

template <class TB>
__global__ void __launch_bounds__(1024) mlp_wgrad_kernel(uint32_t M, uint32_t N, uint32_t K, const float* A,
                                                         uint32_t a_ld, const TB* B, uint32_t b_ld, float* C,
                                                         uint32_t c_ld, float* bias) {
    __shared__ float s_tile[kWgradWaves][16][17];
    __shared__ float s_bias[kWgradWaves][4][16];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t ar = lane & 15u, kg = lane >> 4;
    const uint32_t i0 = blockIdx.y * 16u, j0 = blockIdx.x * 16u;
    const uint32_t ai = i0 + ar, bj = j0 + ar;
    const bool with_bias = bias != nullptr && blockIdx.x == 0u;
    mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    for (uint32_t p0 = wave * 32u; p0 < K; p0 += kWgradWaves * 32u) {
        mlp_bf16x8 a, b;
#pragma unroll
        for (uint32_t e = 0; e < 8u; ++e) {
            const uint32_t p = p0 + kg * 8u + e;
            const float av = (ai < M && p < K) ? A[(size_t)p * a_ld + ai] : 0.0f;
            bsum += av;
            a[e] = (__bf16)av;
            b[e] = (bj < N && p < K) ? as_bf16(B[(size_t)p * b_ld + bj]) : (__bf16)0.0f;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (uint32_t r = 0; r < 4u; ++r) s_tile[wave][kg * 4u + r][ar] = acc[r];
    s_bias[wave][kg][ar] = bsum;
    __syncthreads();
    if (threadIdx.x < 256u) {
        const uint32_t row = threadIdx.x >> 4, col = threadIdx.x & 15u;
        float v = 0.0f;
        for (uint32_t w = 0; w < kWgradWaves; ++w) v += s_tile[w][row][col];
        const uint32_t i = i0 + row, j = j0 + col;
        if (i < M && j < N) C[(size_t)i * c_ld + j] = v;
    } else if (with_bias && threadIdx.x < 272u) {
        const uint32_t row = threadIdx.x - 256u;
        float v = 0.0f;
        for (uint32_t w = 0; w < kWgradWaves; ++w)
            for (uint32_t g = 0; g < 4u; ++g) v += s_bias[w][g][row];
        if (i0 + row < M) bias[i0 + row] = v;
    }
}

extern "C" __global__ void __launch_bounds__(256) mr_mlp_adam_kernel(float* w, __bf16* w_bf16, const float* g,
                                                                     float* m1, float* m2, uint32_t count, float lr,
                                                                     float beta1, float beta2, float eps, float c1,
                                                                     float c2) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float a = beta1 * m1[i] + (1.0f - beta1) * gi;
    const float b = beta2 * m2[i] + (1.0f - beta2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float ahat = a / c1, bhat = b / c2;
    const float wi = w[i] - lr * ahat / (sqrtf(bhat) + eps);
    w[i] = wi;
    if (w_bf16 != nullptr) w_bf16[i] = (__bf16)wi;
}

template <class TA, class TB>
hipError_t gemm(hipStream_t st, uint32_t M, uint32_t N, uint32_t K, const TA* A, uint32_t a_si, uint32_t a_sp,
                const TB* B, uint32_t b_sp, uint32_t b_sj, float* C, uint32_t c_ld, const float* gate) {
    const dim3 grid((N + 15u) / 16u, (M + 15u) / 16u);
    hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB>), grid, dim3(64), 0, st, M, N, K, A, a_si, a_sp, B, b_sp, b_sj, C,
                       c_ld, gate);
    return hipGetLastError();
}

template <class TB>
hipError_t wgrad(hipStream_t st, uint32_t M, uint32_t N, uint32_t K, const float* A, uint32_t a_ld, const TB* B,
                 uint32_t b_ld, float* C, uint32_t c_ld, float* bias) {
    const dim3 grid((N + 15u) / 16u, (M + 15u) / 16u);
    hipLaunchKernelGGL((mlp_wgrad_kernel<TB>), grid, dim3(64u * kWgradWaves), 0, st, M, N, K, A, a_ld, B, b_ld, C,
                       c_ld, bias);
    return hipGetLastError();
}

MlpView view_of(const MRMlpWeightsC& w) {
    MlpView m;
    m.w1 = reinterpret_cast<const __bf16*>(w.w1);
    m.b1 = w.b1;
    m.w2 = reinterpret_cast<const __bf16*>(w.w2);
    m.b2 = w.b2;
    m.w3 = reinterpret_cast<const __bf16*>(w.w3);
    m.b3 = w.b3;
    m.in_shift = w.in_shift;
    m.in_scale = w.in_scale;
    m.out_scale = w.out_scale;
    return m;
}

mr::MlpTrace trace_of(const MRMlpTraceC& t) {
    mr::MlpTrace r;
    r.xq = reinterpret_cast<__bf16*>(t.xq);
    r.z1 = t.z1;
    r.h1 = reinterpret_cast<__bf16*>(t.h1);
    r.z2 = t.z2;
    r.h2 = reinterpret_cast<__bf16*>(t.h2);
    return r;
}

} // namespace

namespace mr {

hipError_t mlp_forward(const MRMlpWeightsC& w, const float* x, uint32_t n, float* y_hat, const MRMlpTraceC* trace,
                       hipStream_t st) {
    const MlpView m = view_of(w);
    MlpTrace t{};
    if (trace != nullptr) t = trace_of(*trace);
    hipLaunchKernelGGL(mr_mlp_forward_kernel, dim3((n + kMlpRows - 1u) / kMlpRows), dim3(64), 0, st, m, x, n, y_hat,
                       t, trace != nullptr ? 1u : 0u);
    return hipGetLastError();
}

uint32_t mlp_loss_scratch_floats(uint32_t n) { return (n * kMlpOut + 255u) / 256u; }

hipError_t mlp_loss(const float* y_hat, const float* t, const float* out_scale, uint32_t n, float delta, float* dy,
                    float* loss, float* scratch, hipStream_t st) {
    const uint32_t blocks = mlp_loss_scratch_floats(n);
    hipLaunchKernelGGL(mr_mlp_mse_kernel, dim3(blocks), dim3(256), 0, st, y_hat, t, out_scale, n, delta, dy,
                       scratch);
    hipLaunchKernelGGL(mr_mlp_sum_kernel, dim3(1), dim3(256), 0, st, scratch, blocks, 1.0f / (float)(n * kMlpOut),
                       loss);
    return hipGetLastError();
}

hipError_t mlp_backward(const MRMlpWeightsC& w, const MRMlpTraceC& trc, const float* dy, uint32_t n, float* dz2,
                        float* dz1, const MRMlpGradsC& g, hipStream_t st) {
    const MlpView m = view_of(w);
    const MlpTrace tr = trace_of(trc);
    hipError_t err;
    // layer 3: dW3 = dy^T h2, db3 = column sums of dy
    if ((err = wgrad<__bf16>(st, kMlpOut, kMlpHidden, n, dy, kMlpOut, tr.h2, kMlpHidden, g.w3, kMlpHidden, g.b3)) !=
        hipSuccess)
        return err;
    const size_t elems = (size_t)n * kMlpHidden;
    hipLaunchKernelGGL(mr_mlp_out_backward_kernel, dim3((uint32_t)((elems + 255u) / 256u)), dim3(256), 0, st, dy,
                       m.w3, tr.z2, n, dz2);
    // layer 2: dW2 = dz2^T h1, db2 = column sums of dz2
    if ((err = wgrad<__bf16>(st, kMlpHidden, kMlpHidden, n, dz2, kMlpHidden, tr.h1, kMlpHidden, g.w2, kMlpHidden,
                             g.b2)) != hipSuccess)
        return err;
    // dz1 = (dz2 W2) * gelu'(z1)
    if ((err = gemm<float, __bf16>(st, n, kMlpHidden, kMlpHidden, dz2, kMlpHidden, 1u, m.w2, kMlpHidden, 1u, dz1,
                                   kMlpHidden, tr.z1)) != hipSuccess)
        return err;
    // layer 1: dW1 = dz1^T xq, db1 = column sums of dz1
    return wgrad<__bf16>(st, kMlpHidden, kMlpIn, n, dz1, kMlpHidden, tr.xq, kMlpIn, g.w1, kMlpIn, g.b1);
}

hipError_t mlp_adam(float* w, uint16_t* w_bf16, const float* g, float* m1, float* m2, uint32_t count, float lr,
                    float beta1, float beta2, float eps, uint32_t step, hipStream_t st) {
    float c1 = 1.0f, c2 = 1.0f;
    for (uint32_t k = 0; k < step; ++k) {
        c1 *= beta1;
        c2 *= beta2;
    }
    c1 = 1.0f - c1;
    c2 = 1.0f - c2;
    hipLaunchKernelGGL(mr_mlp_adam_kernel, dim3((count + 255u) / 256u), dim3(256), 0, st, w,
                       reinterpret_cast<__bf16*>(w_bf16), g, m1, m2,
                       count, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

} // namespace mr
