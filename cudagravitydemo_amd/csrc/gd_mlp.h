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
    /* hardware conversion, round to nearest even */
    return __builtin_bit_cast(unsigned short, (__bf16)f);
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

/* 16x16x32 product from two pairs of 16x16x16 operands: K index k of the
 * 32-wide block is taken as (lane quarter q, slot j) -> j < 4 ? 4q + j :
 * 16 + 4q + j - 4, so the A operand is the concatenation of the two 16x16x16
 * weight tiles (k tiles 2kb, 2kb+1) a lane already holds, and the B operand
 * the concatenation of the two activation tiles.  Same sum, half the
 * instructions. */
typedef __bf16 gd_bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ gd_bf16x8 gd_cat8(gd_bf16x4 a, gd_bf16x4 b) {
    const gd_bf16x4 v0 = a, v1 = b;
    typedef short gd_s16x8 __attribute__((ext_vector_type(8)));
    const gd_s16x8 r = {v0[0], v0[1], v0[2], v0[3], v1[0], v1[1], v1[2], v1[3]};
    return __builtin_bit_cast(gd_bf16x8, r);
}
__device__ __forceinline__ gd_f32x4 gd_mfma32(gd_bf16x4 a0, gd_bf16x4 a1, gd_bf16x4 b0,
                                              gd_bf16x4 b1, gd_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(gd_cat8(a0, a1), gd_cat8(b0, b1), c, 0, 0,
                                                   0);
}

#define GD_GELU_C 0.7978845608028654f
#define GD_GELU_A 0.044715f

/* GELU, tanh form: 0.5 z (1 + tanh u), u = c (z + a z^3).  With
 * 1 + tanh u = 2 - 2 / (1 + e^{2u}) this is z - z / (1 + e^{2u}): one
 * exponential and one reciprocal (both hardware approximations, ~1 ulp),
 * exact limits z (u -> +inf) and 0 (u -> -inf). */
__device__ __forceinline__ float gd_gelu(float z) {
    /* 2u log2(e) = z (K + K a z^2), K = 2 c log2(e) */
    constexpr float K = 2.f * GD_GELU_C * 1.4426950408889634f;
    const float u2 = z * __builtin_fmaf(K * GD_GELU_A, z * z, K);
    const float r = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(u2));
    return __builtin_fmaf(-z, r, z);
}
/* d/dz: with s = 1 / (1 + e^{2u}) (so tanh u = 1 - 2s):
 * (1 - s) + z * 2 s (1 - s) * c (1 + 3 a z^2). */
__device__ __forceinline__ float gd_gelu_grad(float z) {
    const float u2 = (2.f * GD_GELU_C * 1.4426950408889634f) * __builtin_fmaf(GD_GELU_A * z * z, z, z);
    const float s = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(u2));
    const float du = GD_GELU_C * __builtin_fmaf(3.f * GD_GELU_A, z * z, 1.f);
    return (1.f - s) + z * (2.f * s * (1.f - s)) * du;
}

/* Inference only (no activations kept): layer 1 on 16x16x16 tiles, the
 * 64-wide layers on 16x16x32 (see gd_mfma32).  The bias enters as the
 * accumulator's initial value (rows r0..r0+3 of the lane's column). */
__device__ __forceinline__ gd_f32x4 gd_mlp_infer16(const gd_bf16x4* __restrict__ fwd,
                                                   const float* __restrict__ bias, gd_bf16x4 x) {
    const unsigned l = gd_lane();
    const int r0 = 4 * (int)(l / 16);
    auto b4 = [&](int at) {
        const gd_f32x4 v = {bias[at + r0], bias[at + r0 + 1], bias[at + r0 + 2],
                            bias[at + r0 + 3]};
        return v;
    };
    gd_bf16x4 in[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const gd_f32x4 z = gd_mfma(fwd[t * 64 + l], x, b4(16 * t));
        in[t] = gd_pack4(gd_gelu(z[0]), gd_gelu(z[1]), gd_gelu(z[2]), gd_gelu(z[3]));
    }
#pragma unroll
    for (int L = 1; L <= 2; ++L) {
        const gd_bf16x4* W = fwd + (L == 1 ? 4 : 20) * 64;
        gd_bf16x4 nxt[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            gd_f32x4 z = gd_mfma32(W[(t * 4 + 0) * 64 + l], W[(t * 4 + 1) * 64 + l], in[0], in[1],
                                   b4(64 * L + 16 * t));
            z = gd_mfma32(W[(t * 4 + 2) * 64 + l], W[(t * 4 + 3) * 64 + l], in[2], in[3], z);
            nxt[t] = gd_pack4(gd_gelu(z[0]), gd_gelu(z[1]), gd_gelu(z[2]), gd_gelu(z[3]));
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) in[t] = nxt[t];
    }
    const gd_bf16x4* W4 = fwd + 36 * 64;
    gd_f32x4 o = gd_mfma32(W4[0 * 64 + l], W4[1 * 64 + l], in[0], in[1], b4(192));
    o = gd_mfma32(W4[2 * 64 + l], W4[3 * 64 + l], in[2], in[3], o);
    return o;
}

#endif
