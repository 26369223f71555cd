// Warm-start MLP, device side: one wavefront evaluates 16 rows (contacts) of
// the 8 -> 256 -> 256 -> 3 network with BF16 weights and activations on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16) and FP32 accumulation.
//
//   x_hat = (x - shift) * scale                 (FP32, rounded to BF16)
//   z1 = W1 x_hat + b1,  h1 = gelu(z1)          (h1 rounded to BF16)
//   z2 = W2 h1 + b2,     h2 = gelu(z2)          (h2 rounded to BF16)
//   y_hat = W3 h2 + b3                          (FP32; prediction = out_scale * y_hat)
//
// Weights are stored out-major ([out][in], BF16), so the 8 consecutive inputs
// one lane feeds to an MFMA are one 16-byte load. Operand layout of
// v_mfma_f32_16x16x32_bf16 (64 lanes, l = lane): A supplies row (l % 16) and
// inputs 8 * (l / 16) .. + 7; B supplies output column (l % 16) for the same
// inputs; the 16x16 FP32 result holds row 4 * (l / 16) + r, column l % 16 in
// accumulator element r. The same routine serves training (which also keeps
// z and h) and the in-solver inference, so both give the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mr {

constexpr uint32_t kMlpIn = 8u;
constexpr uint32_t kMlpHidden = 256u;
constexpr uint32_t kMlpOut = 3u;
constexpr uint32_t kMlpRows = 16u;                  // rows per wavefront tile
constexpr uint32_t kMlpHStride = kMlpHidden + 8u;   // padded LDS row (bf16): no bank conflicts
constexpr uint32_t kMlpLdsBytes = kMlpRows * kMlpIn * 4u + 2u * kMlpRows * kMlpHStride * 2u;

typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

// Device pointers of one network.
struct MlpView {
    const __bf16* w1;       // [256][8]
    const float* b1;        // [256]
    const __bf16* w2;       // [256][256]
    const float* b2;        // [256]
    const __bf16* w3;       // [3][256]
    const float* b3;        // [3]
    const float* in_shift;  // [8]
    const float* in_scale;  // [8]
    const float* out_scale; // [3]
};

// Optional per-row records of a training forward pass (global, row-major).
struct MlpTrace {
    __bf16* xq;  // [n][8]   normalized input as fed to the matrix cores
    float* z1;   // [n][256]
    __bf16* h1;  // [n][256]
    float* z2;   // [n][256]
    __bf16* h2;  // [n][256]
};

__device__ __forceinline__ float mlp_gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.7071067811865476f)); }

// d gelu / dx = Phi(x) + x * phi(x)
__device__ __forceinline__ float mlp_gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.7071067811865476f));
    const float pdf = 0.3989422804014327f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// LDS scratch of one tile: normalized inputs and the two hidden layers.
struct MlpLds {
    float* x;    // [16][8] FP32 normalized inputs
    __bf16* h1;  // [16][kMlpHStride]
    __bf16* h2;  // [16][kMlpHStride]
};

__device__ __forceinline__ MlpLds mlp_lds(unsigned char* base) {
    MlpLds s;
    s.x = reinterpret_cast<float*>(base);
    s.h1 = reinterpret_cast<__bf16*>(base + kMlpRows * kMlpIn * 4u);
    s.h2 = s.h1 + kMlpRows * kMlpHStride;
    return s;
}

// One hidden layer from an LDS source (K = 256) into an LDS destination.
__device__ __forceinline__ void mlp_hidden_layer(const __bf16* src, const __bf16* w, const float* b, __bf16* dst,
                                                 uint32_t lane, float* z_out, __bf16* h_out, uint32_t row0,
                                                 uint32_t rows) {
    const uint32_t ar = lane & 15u, kg = lane >> 4;
    for (uint32_t nb = 0; nb < kMlpHidden / 16u; ++nb) {
        mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint32_t col = nb * 16u + ar;
#pragma unroll
        for (uint32_t kb = 0; kb < kMlpHidden / 32u; ++kb) {
            const uint32_t k0 = kb * 32u + kg * 8u;
            const mlp_bf16x8 a = *reinterpret_cast<const mlp_bf16x8*>(src + ar * kMlpHStride + k0);
            const mlp_bf16x8 bv = *reinterpret_cast<const mlp_bf16x8*>(w + (size_t)col * kMlpHidden + k0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc, 0, 0, 0);
        }
        const float bias = b[col];
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) {
            const uint32_t row = kg * 4u + r;
            const float z = acc[r] + bias;
            const __bf16 h = (__bf16)mlp_gelu(z);
            dst[row * kMlpHStride + col] = h;
            if (z_out != nullptr && row < rows) {
                z_out[(size_t)(row0 + row) * kMlpHidden + col] = z;
                h_out[(size_t)(row0 + row) * kMlpHidden + col] = h;
            }
        }
    }
}

// Evaluates rows row0 .. row0 + 15 (those below `rows` are real; the rest
// carry zero inputs). `s.x` must already hold the normalized inputs of the
// tile. Lanes whose column (lane % 16) is below 3 return y_hat of rows
// 4 * (lane / 16) + r in y[r]. All 64 lanes must be active.
__device__ __forceinline__ void mlp_tile_forward(const MlpView& m, const MlpLds& s, uint32_t lane, float y[4],
                                                 const MlpTrace* trace, uint32_t row0, uint32_t rows) {
    const uint32_t ar = lane & 15u, kg = lane >> 4;
    if (trace != nullptr && trace->xq != nullptr) {
        for (uint32_t idx = lane; idx < kMlpRows * kMlpIn; idx += 64u) {
            const uint32_t row = idx / kMlpIn, f = idx % kMlpIn;
            if (row < rows) trace->xq[(size_t)(row0 + row) * kMlpIn + f] = (__bf16)s.x[idx];
        }
    }
    // layer 1: K = 8, carried by lanes 0-15 (the other K groups are zero)
    for (uint32_t nb = 0; nb < kMlpHidden / 16u; ++nb) {
        mlp_bf16x8 a, bv;
        const uint32_t col = nb * 16u + ar;
#pragma unroll
        for (uint32_t e = 0; e < 8u; ++e) {
            a[e] = kg == 0u ? (__bf16)s.x[ar * kMlpIn + e] : (__bf16)0.0f;
            bv[e] = kg == 0u ? m.w1[(size_t)col * kMlpIn + e] : (__bf16)0.0f;
        }
        mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc, 0, 0, 0);
        const float bias = m.b1[col];
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) {
            const uint32_t row = kg * 4u + r;
            const float z = acc[r] + bias;
            const __bf16 h = (__bf16)mlp_gelu(z);
            s.h1[row * kMlpHStride + col] = h;
            if (trace != nullptr && row < rows) {
                trace->z1[(size_t)(row0 + row) * kMlpHidden + col] = z;
                trace->h1[(size_t)(row0 + row) * kMlpHidden + col] = h;
            }
        }
    }
    __syncthreads();
    mlp_hidden_layer(s.h1, m.w2, m.b2, s.h2, lane, trace != nullptr ? trace->z2 : nullptr,
                     trace != nullptr ? trace->h2 : nullptr, row0, rows);
    __syncthreads();
    // layer 3: N = 3, columns 3..15 of the tile are zero
    mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (uint32_t kb = 0; kb < kMlpHidden / 32u; ++kb) {
        const uint32_t k0 = kb * 32u + kg * 8u;
        const mlp_bf16x8 a = *reinterpret_cast<const mlp_bf16x8*>(s.h2 + ar * kMlpHStride + k0);
        mlp_bf16x8 bv;
        if (ar < kMlpOut) {
            bv = *reinterpret_cast<const mlp_bf16x8*>(m.w3 + (size_t)ar * kMlpHidden + k0);
        } else {
#pragma unroll
            for (uint32_t e = 0; e < 8u; ++e) bv[e] = (__bf16)0.0f;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc, 0, 0, 0);
    }
    const float bias = ar < kMlpOut ? m.b3[ar] : 0.0f;
#pragma unroll
    for (uint32_t r = 0; r < 4u; ++r) y[r] = acc[r] + bias;
}

} // namespace mr
