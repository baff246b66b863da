// The warm-start surrogate: a 3 -> 64 -> 64 -> 1 multilayer perceptron with
// tanh-form GELU activations.
//
// Inputs (normalised with the stored mean/scale): ln p2, the parcel's
// equivalent potential temperature (the reference's own parcel theta-e,
// getthe(p, t, t, qv), CAPE_CODE_model_lev.f90:505), and its total water
// mixing ratio.  Output: the sub-step's potential-temperature rate
// r = (theta2 - theta1) / ln(p2/p1), de-normalised with the stored scale and
// offset, from which the first iterate is theta1 + r ln(p2/p1).
//
// Precision: weights are BF16 (W1, w3 as BF16-rounded floats on the VALU,
// W2 as BF16 MFMA operands); the first hidden layer is rounded to BF16 where
// it feeds the MFMA; all accumulation is FP32.
//
// Parameter vector (FP32 master copy, length MLP_NPARAMS): W1 (64x3), b1,
// W2 (64x64, W2[i][j] maps hidden-1 unit j to hidden-2 unit i), b2, w3, b3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xcape {

constexpr int MLP_H = 64;
constexpr int MLP_IN = 3;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NPARAMS = MLP_OFF_B3 + 1;   // 4481

// Device copy the ascent kernels read (built from the FP32 master vector).
struct SurrogateDev {
    uint16_t w2[MLP_H * MLP_H];   // BF16 bits, row-major [i][j]
    float w1[MLP_H * MLP_IN];     // BF16-rounded
    float b1[MLP_H];
    float b2[MLP_H];
    float w3[MLP_H];              // BF16-rounded
    float b3;
    float in_mean[MLP_IN], in_scale[MLP_IN];
    float out_scale, out_offset;
};

// round to nearest even to BF16 (bits / value); NaN stays NaN
__host__ __device__ __forceinline__ uint16_t bf16_bits(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_round(float f)
{
    uint32_t u = (uint32_t)bf16_bits(f) << 16;
    float r;
    __builtin_memcpy(&r, &u, 4);
    return r;
}

// tanh-form GELU and its derivative, FP32: 0.5 z (1 + tanh(c (z + a z^3))).
// tanh(u) = 1 - 2 / (1 + exp(2u)), with the hardware exp and reciprocal
// (about 1 ulp, far below the BF16 rounding of the activations); saturates
// cleanly for large |u| (the reciprocal of inf is 0).
__device__ __forceinline__ float gelu_f(float z)
{
    const float c = 0.7978845608028654f, a = 0.044715f;
    const float u = c * (z + a * z * z * z);
    const float e = __expf(2.0f * u);
    const float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
    return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad_f(float z)
{
    const float c = 0.7978845608028654f, a = 0.044715f;
    const float u = c * (z + a * z * z * z);
    const float e = __expf(2.0f * u);
    const float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * a * z * z);
}

typedef __bf16 xc_bf16x8 __attribute__((ext_vector_type(8)));
typedef float xc_f32x16 __attribute__((ext_vector_type(16)));

// LDS the wave needs for one evaluation: the BF16 hidden-1 rows of 32
// samples (one half of the wave at a time), padded to 33 words per row so
// that 32 lanes writing their own rows hit distinct banks.  4.2 KB per wave,
// which keeps 8 waves per SIMD within the CU's LDS when the evaluation
// shares its kernel with the ascent.
constexpr int MLP_LDS_ROW_WORDS = 33;
constexpr int MLP_LDS_WORDS_PER_WAVE = 32 * MLP_LDS_ROW_WORDS;

// Evaluate the network for the 64 samples of a wavefront (one per lane).
// Must be called by every lane of the block together (the MFMA uses all 64
// lanes of a wave and the LDS hand-over uses block barriers); `lds` points
// to this wave's MLP_LDS_WORDS_PER_WAVE words.
//
// Layer 2 runs as v_mfma_f32_32x32x16_bf16: Z2^T (hidden-2 x samples) =
// W2 (hidden-2 x hidden-1) . H1^T (hidden-1 x samples), four 32x32 output
// tiles (tm: hidden-2 rows, tn: samples), each accumulated over 4 K-steps of
// 16.  Operand layout per lane l: A[m = l%32][k = 8(l/32) .. +7],
// B[k = 8(l/32) .. +7][n = l%32]; result register r of lane l is
// C[8(r/4) + 4(l/32) + r%4][l%32].  Phase tn stages the hidden-1 rows of
// samples 32 tn .. 32 tn + 31 (computed by the lanes of half tn).  Tiles and
// K-steps are not unrolled, so the live set is one accumulator and two
// operands (the ascent that follows is register-bound).
__device__ __forceinline__ float surrogate_eval_wave(const SurrogateDev *S, float x0, float x1,
                                                     float x2, uint32_t *lds)
{
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5, col = lane & 31;
    const float u0 = (x0 - S->in_mean[0]) * S->in_scale[0];
    const float u1 = (x1 - S->in_mean[1]) * S->in_scale[1];
    const float u2 = (x2 - S->in_mean[2]) * S->in_scale[2];
    float out0 = 0.0f, out1 = 0.0f;
#pragma unroll 1
    for (int tn = 0; tn < 2; ++tn) {
        // ---- layer 1 on the VALU: the hidden-1 rows of samples 32 tn ..
        // 32 tn + 31; lane (half, col) computes units 32 half .. 32 half + 31
        // of sample 32 tn + col, so that every lane works.  The weights of a
        // unit are wave-uniform loads, selected per half ----
        const int src = 32 * tn + col;
        const float v0 = __shfl(u0, src, 64), v1 = __shfl(u1, src, 64), v2 = __shfl(u2, src, 64);
        uint32_t *row = lds + col * MLP_LDS_ROW_WORDS + 16 * half;
#pragma unroll 4
        for (int j = 0; j < MLP_H / 2; j += 2) {
            const int k = j + MLP_H / 2;
            const float za = half ? S->b1[k] + S->w1[3 * k] * v0 + S->w1[3 * k + 1] * v1 +
                                        S->w1[3 * k + 2] * v2
                                  : S->b1[j] + S->w1[3 * j] * v0 + S->w1[3 * j + 1] * v1 +
                                        S->w1[3 * j + 2] * v2;
            const float zb = half ? S->b1[k + 1] + S->w1[3 * k + 3] * v0 +
                                        S->w1[3 * k + 4] * v1 + S->w1[3 * k + 5] * v2
                                  : S->b1[j + 1] + S->w1[3 * j + 3] * v0 +
                                        S->w1[3 * j + 4] * v1 + S->w1[3 * j + 5] * v2;
            row[j >> 1] = (uint32_t)bf16_bits(gelu_f(za)) |
                          ((uint32_t)bf16_bits(gelu_f(zb)) << 16);
        }
        __syncthreads();
        // ---- layer 2 (MFMA) + layer 3 for the two tiles of these samples ----
#pragma unroll 1
        for (int tm = 0; tm < 2; ++tm) {
            xc_f32x16 acc = {};
#pragma unroll 1
            for (int kb = 0; kb < 4; ++kb) {
                const int k0 = 16 * kb + 8 * half;
                // A: W2[32 tm + col][k0 .. k0+7]; B: H1[sample 32 tn + col][k0 .. k0+7]
                const uint4 wq =
                    *reinterpret_cast<const uint4 *>(S->w2 + (32 * tm + col) * MLP_H + k0);
                const uint32_t *hb = lds + col * MLP_LDS_ROW_WORDS + (k0 >> 1);
                const uint4 hq = make_uint4(hb[0], hb[1], hb[2], hb[3]);
                xc_bf16x8 a, b;
                __builtin_memcpy(&a, &wq, 16);
                __builtin_memcpy(&b, &hq, 16);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
            float part = 0.0f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = 32 * tm + 8 * (r >> 2) + 4 * half + (r & 3);
                part += S->w3[i] * gelu_f(acc[r] + S->b2[i]);
            }
            if (tn == 0) out0 += part; else out1 += part;
        }
        __syncthreads();   // the rows are rewritten by the next phase / caller
    }
    out0 += __shfl_xor(out0, 32, 64);
    out1 += __shfl_xor(out1, 32, 64);
    const float y = (half ? out1 : out0) + S->b3;
    return y * S->out_scale + S->out_offset;
}

}  // namespace xcape
