/*
 * gd_mlp.h -- corrector network constants and operand types shared by the
 * training kernels and the walk.
 *
 * Parameter vector (FP32, row-major [out][in]):
 *   W1[64][8] @0, b1[64] @512, W2[64][64] @576, b2 @4672, W3[64][64] @4736,
 *   b3 @8832, W4[3][64] @8896, b4[3] @9088; 9091 values.
 */
#ifndef GD_MLP_H
#define GD_MLP_H

#include "gd_internal.h"

#define GD_MLP_NPARAMS 9091
#define GD_MLP_W1 0
#define GD_MLP_B1 512
#define GD_MLP_W2 576
#define GD_MLP_B2 4672
#define GD_MLP_W3 4736
#define GD_MLP_B3 8832
#define GD_MLP_W4 8896
#define GD_MLP_B4 9088
#define GD_MLP_FWD_TILES 40   /* W1 4x1, W2 4x4, W3 4x4, W4 1x4 (16x16 BF16 tiles) */
#define GD_MLP_BWD_TILES 36   /* W2^T 4x4, W3^T 4x4, W4^T 4x1 */
#define GD_MLP_BIAS_PAD 208   /* b1 | b2 | b3 | b4 padded to 16 */
#define GD_MLP_PARTIAL (40 * 256 + GD_MLP_BIAS_PAD)

typedef short gd_bf16x4 __attribute__((ext_vector_type(4)));
typedef float gd_f32x4 __attribute__((ext_vector_type(4)));

struct gd_mlp;
const gd_bf16x4* gd_mlp_fwd_tiles(const gd_mlp* net);
const float* gd_mlp_bias(const gd_mlp* net);

/* ---------------------------------------------------------------------------
 * BF16 helpers
 * ------------------------------------------------------------------------- */
__device__ __forceinline__ unsigned short gd_f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);   /* round to nearest even */
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float gd_bf2f(unsigned short h) {
    return __uint_as_float((uint32_t)h << 16);
}
__device__ __forceinline__ gd_bf16x4 gd_pack4(float a, float b, float c, float d) {
    gd_bf16x4 r;
    r[0] = (short)gd_f2bf(a);
    r[1] = (short)gd_f2bf(b);
    r[2] = (short)gd_f2bf(c);
    r[3] = (short)gd_f2bf(d);
    return r;
}
__device__ __forceinline__ gd_f32x4 gd_mfma(gd_bf16x4 a, gd_bf16x4 b, gd_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

#define GD_GELU_C 0.7978845608028654f
#define GD_GELU_A 0.044715f

__device__ __forceinline__ float gd_gelu(float z) {
    const float u = GD_GELU_C * __builtin_fmaf(GD_GELU_A * z * z, z, z);
    const float t = 1.f - 2.f / (__expf(2.f * u) + 1.f);
    return 0.5f * z * (1.f + t);
}
__device__ __forceinline__ float gd_gelu_grad(float z) {
    const float u = GD_GELU_C * __builtin_fmaf(GD_GELU_A * z * z, z, z);
    const float t = 1.f - 2.f / (__expf(2.f * u) + 1.f);
    return 0.5f * (1.f + t) +
           0.5f * z * (1.f - t * t) * GD_GELU_C * __builtin_fmaf(3.f * GD_GELU_A, z * z, 1.f);
}

/* Inference only: the same arithmetic as gd_mlp_forward16, no activations kept. */
__device__ __forceinline__ gd_f32x4 gd_mlp_infer16(const gd_bf16x4* __restrict__ fwd,
                                                   const float* __restrict__ bias, gd_bf16x4 x) {
    const unsigned l = gd_lane();
    const int r0 = 4 * (int)(l / 16);
    gd_f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    gd_bf16x4 in[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        gd_f32x4 z = gd_mfma(fwd[t * 64 + l], x, zero);
#pragma unroll
        for (int j = 0; j < 4; ++j) z[j] += bias[16 * t + r0 + j];
        in[t] = gd_pack4(gd_gelu(z[0]), gd_gelu(z[1]), gd_gelu(z[2]), gd_gelu(z[3]));
    }
#pragma unroll
    for (int L = 1; L <= 2; ++L) {
        const gd_bf16x4* W = fwd + (L == 1 ? 4 : 20) * 64;
        gd_bf16x4 nxt[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            gd_f32x4 z = zero;
#pragma unroll
            for (int k = 0; k < 4; ++k) z = gd_mfma(W[(t * 4 + k) * 64 + l], in[k], z);
#pragma unroll
            for (int j = 0; j < 4; ++j) z[j] += bias[64 * L + 16 * t + r0 + j];
            nxt[t] = gd_pack4(gd_gelu(z[0]), gd_gelu(z[1]), gd_gelu(z[2]), gd_gelu(z[3]));
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) in[t] = nxt[t];
    }
    gd_f32x4 o = zero;
    const gd_bf16x4* W4 = fwd + 36 * 64;
#pragma unroll
    for (int k = 0; k < 4; ++k) o = gd_mfma(W4[k * 64 + l], in[k], o);
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] += bias[192 + r0 + j];
    return o;
}

#endif
