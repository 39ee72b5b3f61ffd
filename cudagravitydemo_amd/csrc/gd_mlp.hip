/*
 * gd_mlp.hip -- the far-field corrector network on the matrix cores.
 *
 * Network 8 -> 64 -> 64 -> 64 -> 3, GELU (tanh form) after the hidden layers.
 * Parameters: one FP32 master vector (layout in gravity_ext.h), Adam moments,
 * and BF16 copies packed as MFMA operand tiles.
 *
 * Every product is v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32
 * accumulation).  The layers are evaluated transposed, Z^T = W H^T: the
 * 16x16 D tile of one layer (lane l holds rows 4(l/16)..4(l/16)+3 of column
 * l%16) is exactly the B operand the next layer needs (lane l holds k rows
 * 4(l/16)..+3 of column l%16), so activations stay in registers between
 * layers; a column is one sample (or, in the walk, one cell).
 *
 * Training (k_mlp_grad): one wavefront per block walks a slice of the batch
 * in chunks of 16 samples: forward, loss gradient, back-propagation of the
 * deltas with W^T tiles on the matrix cores, and weight gradients
 * dW = dZ^T H accumulated with MFMA over the sample dimension (the D-layout
 * registers are transposed through LDS to become A/B operands).  Each block
 * writes its partial gradient; k_mlp_adam sums the partials in a fixed order
 * and applies Adam to the FP32 master, and k_mlp_pack re-packs the BF16 tiles.
 */
#include "gd_internal.h"
#include "gd_mlp.h"
#include "gd_traverse.h"

#include <math.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * Packing: FP32 master -> BF16 operand tiles
 *   forward tiles (A = W, [out][in]):  W1 4x1, W2 4x4, W3 4x4, W4 1x4
 *   backward tiles (A = W^T, [in][out]):          W2 4x4, W3 4x4, W4 4x1
 * tile element for lane l, j: row 16 rt + l%16, col 16 kt + 4 (l/16) + j
 * ------------------------------------------------------------------------- */
__device__ __forceinline__ float gd_w(const float* __restrict__ p, int layer, int o, int i) {
    /* parameter-vector offset of W_layer[o][i], zero outside the real shape */
    const int outs[4] = {64, 64, 64, 3};
    const int ins[4] = {8, 64, 64, 64};
    const int base[4] = {GD_MLP_W1, GD_MLP_W2, GD_MLP_W3, GD_MLP_W4};
    if (o >= outs[layer] || i >= ins[layer]) return 0.f;
    return p[base[layer] + o * ins[layer] + i];
}

__global__ __launch_bounds__(64) void k_mlp_pack(const float* __restrict__ master,
                                                 gd_bf16x4* __restrict__ fwd,
                                                 gd_bf16x4* __restrict__ bwd,
                                                 float* __restrict__ bias) {
    const int tile = blockIdx.x;   /* 0..75 */
    const int l = threadIdx.x;
    const int r = l % 16, c0 = 4 * (l / 16);
    int layer, rt, kt;
    bool transposed;
    if (tile < GD_MLP_FWD_TILES) {
        transposed = false;
        if (tile < 4) { layer = 0; rt = tile; kt = 0; }
        else if (tile < 20) { layer = 1; rt = (tile - 4) / 4; kt = (tile - 4) % 4; }
        else if (tile < 36) { layer = 2; rt = (tile - 20) / 4; kt = (tile - 20) % 4; }
        else { layer = 3; rt = 0; kt = tile - 36; }
    } else {
        const int t = tile - GD_MLP_FWD_TILES;
        transposed = true;
        if (t < 16) { layer = 1; rt = t / 4; kt = t % 4; }
        else if (t < 32) { layer = 2; rt = (t - 16) / 4; kt = (t - 16) % 4; }
        else { layer = 3; rt = t - 32; kt = 0; }
    }
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = 16 * rt + r, col = 16 * kt + c0 + j;
        v[j] = transposed ? gd_w(master, layer, col, row) : gd_w(master, layer, row, col);
    }
    const gd_bf16x4 packed = gd_pack4(v[0], v[1], v[2], v[3]);
    if (transposed) bwd[(tile - GD_MLP_FWD_TILES) * 64 + l] = packed;
    else fwd[tile * 64 + l] = packed;
    if (tile == 0) {
        /* biases, zero-padded: b1 [0,64) b2 [64,128) b3 [128,192) b4 [192,208) */
        for (int k = l; k < GD_MLP_BIAS_PAD; k += 64) {
            float b = 0.f;
            if (k < 64) b = master[GD_MLP_B1 + k];
            else if (k < 128) b = master[GD_MLP_B2 + k - 64];
            else if (k < 192) b = master[GD_MLP_B3 + k - 128];
            else if (k < 195) b = master[GD_MLP_B4 + k - 192];
            bias[k] = b;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Forward pass of 16 columns held in B-operand layout (x: features
 * 4(l/16)..+3 of column l%16, features >= 8 zero).  Hidden activations are
 * returned in B layout (BF16) with their pre-activations; out[j] is output
 * 4(l/16)+j of column l%16 (valid for l < 16, j < 3).
 * ------------------------------------------------------------------------- */
struct gd_mlp_acts {
    gd_f32x4 z[3][4];      /* pre-activations of the hidden layers, D layout */
    gd_bf16x4 h[3][4];     /* GELU(z) as BF16 B operands */
};

__device__ __forceinline__ gd_f32x4 gd_mlp_forward16(const gd_bf16x4* __restrict__ fwd,
                                                     const float* __restrict__ bias,
                                                     gd_bf16x4 x, gd_mlp_acts* acts) {
    const unsigned l = gd_lane();
    const int r0 = 4 * (int)(l / 16);
    gd_f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    gd_bf16x4 in[4];
    /* layer 1: 4 row tiles, one k tile */
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        gd_f32x4 z = gd_mfma(fwd[t * 64 + l], x, zero);
#pragma unroll
        for (int j = 0; j < 4; ++j) z[j] += bias[16 * t + r0 + j];
        acts->z[0][t] = z;
        in[t] = gd_pack4(gd_gelu(z[0]), gd_gelu(z[1]), gd_gelu(z[2]), gd_gelu(z[3]));
        acts->h[0][t] = in[t];
    }
    /* layers 2 and 3 */
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
            acts->z[L][t] = z;
            nxt[t] = gd_pack4(gd_gelu(z[0]), gd_gelu(z[1]), gd_gelu(z[2]), gd_gelu(z[3]));
            acts->h[L][t] = nxt[t];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) in[t] = nxt[t];
    }
    /* layer 4: one row tile (3 real rows), 4 k tiles */
    gd_f32x4 o = zero;
    const gd_bf16x4* W4 = fwd + 36 * 64;
#pragma unroll
    for (int k = 0; k < 4; ++k) o = gd_mfma(W4[k * 64 + l], in[k], o);
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] += bias[192 + r0 + j];
    return o;
}

/* Load 16 feature rows (row-major, 8 floats each) as a B operand. */
__device__ __forceinline__ gd_bf16x4 gd_load_features(const float* __restrict__ X, long long row0,
                                                      long long rows) {
    const unsigned l = gd_lane();
    const long long s = row0 + (l % 16);
    const int f0 = 4 * (int)(l / 16);
    if (s >= rows || f0 >= 8) return gd_pack4(0.f, 0.f, 0.f, 0.f);
    const float4 v = *reinterpret_cast<const float4*>(X + s * 8 + f0);
    return gd_pack4(v.x, v.y, v.z, v.w);
}

__global__ __launch_bounds__(64) void k_mlp_infer(const gd_bf16x4* __restrict__ fwd,
                                                  const float* __restrict__ bias,
                                                  const float* __restrict__ X, float* __restrict__ Y,
                                                  long long rows) {
    const unsigned l = gd_lane();
    for (long long row0 = (long long)blockIdx.x * 16; row0 < rows; row0 += (long long)gridDim.x * 16) {
        const gd_bf16x4 x = gd_load_features(X, row0, rows);
        const gd_f32x4 o = gd_mlp_infer16(fwd, bias, x);
        const long long s = row0 + (l % 16);
        if (l < 16 && s < rows) {
            Y[s * 3 + 0] = o[0];
            Y[s * 3 + 1] = o[1];
            Y[s * 3 + 2] = o[2];
        }
    }
}

/* ---------------------------------------------------------------------------
 * Training: gradients
 * ------------------------------------------------------------------------- */

/* Transpose a D-layout 16x16 tile (rows = features, cols = samples) through
 * LDS into A layout (rows = features, k = samples) or B layout (k = samples,
 * cols = features) -- both read "feature f, sample s" with f = l%16 or
 * s = l%16 respectively. */
__device__ __forceinline__ void gd_tile_store(float* sm, gd_f32x4 v) {
    const unsigned l = gd_lane();
#pragma unroll
    for (int j = 0; j < 4; ++j) sm[(4 * (l / 16) + j) * 17 + (l % 16)] = v[j];   /* [feature][sample] */
}
/* A operand of dW = dZ^T H: A[i = out feature l%16][k = sample 4(l/16)+j] */
__device__ __forceinline__ gd_bf16x4 gd_tile_as_rows(const float* sm) {
    const unsigned l = gd_lane();
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = sm[(l % 16) * 17 + 4 * (l / 16) + j];
    return gd_pack4(v[0], v[1], v[2], v[3]);
}
/* B operand of dW = dZ^T H: B[k = sample 4(l/16)+j][j' = in feature l%16] */
__device__ __forceinline__ gd_bf16x4 gd_tile_as_cols(const float* sm) {
    return gd_tile_as_rows(sm);   /* same access: element (feature l%16, sample 4(l/16)+j) */
}

#define GD_MLP_GRAD_TILES 40   /* dW tiles: W1 4x1, W2 4x4, W3 4x4, W4 1x4 */

/* accumulate one dW tile in LDS: C = LDS tile, D = A B + C */
__device__ __forceinline__ void gd_acc_tile(gd_f32x4* __restrict__ acc, gd_bf16x4 a, gd_bf16x4 b) {
    const unsigned l = gd_lane();
    acc[l] = gd_mfma(a, b, acc[l]);
}

/* add the 16 sample columns of a D-layout value into bias accumulator slots */
__device__ __forceinline__ void gd_acc_bias(float* __restrict__ gb, int base, gd_f32x4 v) {
    const unsigned l = gd_lane();
    const int r0 = 4 * (int)(l / 16);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float x = v[j];
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) x += __shfl_xor(x, off, 16);
        if ((l % 16) == 0) gb[base + r0 + j] += x;
    }
}

__global__ __launch_bounds__(64) void k_mlp_grad(const gd_bf16x4* __restrict__ fwd,
                                                 const gd_bf16x4* __restrict__ bwd,
                                                 const float* __restrict__ bias,
                                                 const float* __restrict__ X,
                                                 const float* __restrict__ Y,
                                                 const int* __restrict__ order, long long B,
                                                 int perBlock, float* __restrict__ partial,
                                                 float* __restrict__ lossPart) {
    __shared__ gd_f32x4 sGW[GD_MLP_GRAD_TILES][64];   /* dW tiles, D layout */
    __shared__ float sGB[GD_MLP_BIAS_PAD];
    __shared__ float sT[2][16 * 17];                   /* transpose scratch */
    __shared__ float sX[16 * 9];
    const unsigned l = gd_lane();
    const int r0 = 4 * (int)(l / 16);
    gd_f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < GD_MLP_GRAD_TILES; ++t) sGW[t][l] = zero;
    for (int k = l; k < GD_MLP_BIAS_PAD; k += 64) sGB[k] = 0.f;
    float loss = 0.f;
    const float inv = 2.f / (float)B;
    __syncthreads();

    const long long begin = (long long)blockIdx.x * perBlock;
    long long end = begin + perBlock;
    if (end > B) end = B;
    for (long long row0 = begin; row0 < end; row0 += 16) {
        /* the chunk's 16 samples (through `order`), row-major in LDS */
        for (int q = l; q < 16 * 8; q += 64) {
            const long long s = row0 + q / 8;
            float v = 0.f;
            if (s < end) v = X[(long long)order[s] * 8 + (q % 8)];
            sX[(q / 8) * 9 + (q % 8)] = v;
        }
        __syncthreads();
        gd_bf16x4 x;
        {
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = (r0 + j) < 8 ? sX[(l % 16) * 9 + r0 + j] : 0.f;
            x = gd_pack4(v[0], v[1], v[2], v[3]);
        }
        gd_mlp_acts a;
        const gd_f32x4 o = gd_mlp_forward16(fwd, bias, x, &a);
        /* output delta: 2 (y_hat - y) / B for real outputs and samples */
        const long long s = row0 + (l % 16);
        gd_f32x4 d4 = zero;
        if (s < end) {
            const long long row = order[s];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (r0 + j < 3) {
                    const float diff = o[j] - Y[row * 3 + r0 + j];
                    loss += diff * diff;
                    d4[j] = inv * diff;
                }
            }
        }
        gd_acc_bias(sGB, 192, d4);
        /* dW4 (16 x 64) += d4^T (16 x 16 samples) . H3 (16 samples x 64) */
        __syncthreads();
        gd_tile_store(sT[0], d4);
        __syncthreads();
        const gd_bf16x4 d4rows = gd_tile_as_rows(sT[0]);
        const gd_bf16x4 d4b = gd_pack4(d4[0], d4[1], d4[2], d4[3]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gd_f32x4 hv;
#pragma unroll
            for (int j = 0; j < 4; ++j) hv[j] = gd_bf2f((unsigned short)a.h[2][k][j]);
            __syncthreads();
            gd_tile_store(sT[1], hv);
            __syncthreads();
            gd_acc_tile(sGW[36 + k], d4rows, gd_tile_as_cols(sT[1]));
        }
        /* back through the hidden layers: dH = W^T dZ, dZ = dH * gelu'(z) */
        gd_f32x4 dz[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const gd_f32x4 dh = gd_mfma(bwd[(32 + t) * 64 + l], d4b, zero);
#pragma unroll
            for (int j = 0; j < 4; ++j) dz[t][j] = dh[j] * gd_gelu_grad(a.z[2][t][j]);
        }
#pragma unroll
        for (int L = 2; L >= 0; --L) {
            const int baseTile = L == 0 ? 0 : (L == 1 ? 4 : 20);
            const int kTiles = L == 0 ? 1 : 4;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                gd_acc_bias(sGB, 64 * L + 16 * t, dz[t]);
                __syncthreads();
                gd_tile_store(sT[0], dz[t]);
                __syncthreads();
                const gd_bf16x4 rows = gd_tile_as_rows(sT[0]);
#pragma unroll
                for (int k = 0; k < kTiles; ++k) {
                    gd_f32x4 hv;
                    if (L == 0) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) hv[j] = gd_bf2f((unsigned short)x[j]);
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            hv[j] = gd_bf2f((unsigned short)a.h[L - 1][k][j]);
                    }
                    __syncthreads();
                    gd_tile_store(sT[1], hv);
                    __syncthreads();
                    gd_acc_tile(sGW[baseTile + t * kTiles + k], rows, gd_tile_as_cols(sT[1]));
                }
            }
            if (L == 0) break;
            gd_bf16x4 dnext[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) dnext[t] = gd_pack4(dz[t][0], dz[t][1], dz[t][2], dz[t][3]);
            const gd_bf16x4* WT = bwd + (L == 1 ? 0 : 16) * 64;
            gd_f32x4 nz[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                gd_f32x4 dh = zero;
#pragma unroll
                for (int k = 0; k < 4; ++k) dh = gd_mfma(WT[(t * 4 + k) * 64 + l], dnext[k], dh);
#pragma unroll
                for (int j = 0; j < 4; ++j) nz[t][j] = dh[j] * gd_gelu_grad(a.z[L - 1][t][j]);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) dz[t] = nz[t];
        }
        __syncthreads();
    }
    /* this block's partial: the 40 dW tiles in D layout, then the biases */
    float* out = partial + (size_t)blockIdx.x * GD_MLP_PARTIAL;
    for (int t = 0; t < GD_MLP_GRAD_TILES; ++t) {
        const gd_f32x4 v = sGW[t][l];
#pragma unroll
        for (int j = 0; j < 4; ++j) out[(t * 64 + l) * 4 + j] = v[j];
    }
    for (int k = l; k < GD_MLP_BIAS_PAD; k += 64) out[GD_MLP_GRAD_TILES * 256 + k] = sGB[k];
    float ls = loss;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off);
    if (l == 0) lossPart[blockIdx.x] = ls;
}

/* Map a parameter index to its slot in a block partial (tile, lane, j). */
__device__ __forceinline__ int gd_grad_slot(int p) {
    int layer, o, i, ins;
    if (p < GD_MLP_B1) { layer = 0; ins = 8; o = p / ins; i = p % ins; }
    else if (p < GD_MLP_W2) return GD_MLP_GRAD_TILES * 256 + (p - GD_MLP_B1);
    else if (p < GD_MLP_B2) { layer = 1; ins = 64; o = (p - GD_MLP_W2) / ins; i = (p - GD_MLP_W2) % ins; }
    else if (p < GD_MLP_W3) return GD_MLP_GRAD_TILES * 256 + 64 + (p - GD_MLP_B2);
    else if (p < GD_MLP_B3) { layer = 2; ins = 64; o = (p - GD_MLP_W3) / ins; i = (p - GD_MLP_W3) % ins; }
    else if (p < GD_MLP_W4) return GD_MLP_GRAD_TILES * 256 + 128 + (p - GD_MLP_B3);
    else if (p < GD_MLP_B4) { layer = 3; ins = 64; o = (p - GD_MLP_W4) / ins; i = (p - GD_MLP_W4) % ins; }
    else return GD_MLP_GRAD_TILES * 256 + 192 + (p - GD_MLP_B4);
    const int kTiles = layer == 0 ? 1 : 4;
    const int base = layer == 0 ? 0 : (layer == 1 ? 4 : (layer == 2 ? 20 : 36));
    const int t = o / 16, k = i / 16;
    const int tile = base + t * kTiles + k;
    /* D layout: lane = column (i % 16) + 16 * ((o % 16) / 4), j = (o % 16) % 4 */
    const int lane = (i % 16) + 16 * ((o % 16) / 4);
    const int j = (o % 16) % 4;
    return (tile * 64 + lane) * 4 + j;
}

/* Sum the block partials (fixed order) into grad[p]; optionally Adam-step. */
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(const float* __restrict__ partial,
                                                         int blocks, float* __restrict__ grad,
                                                         float* __restrict__ master,
                                                         float* __restrict__ m1,
                                                         float* __restrict__ m2, int doAdam,
                                                         float lr, float beta1, float beta2,
                                                         float eps, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= GD_MLP_NPARAMS) return;
    const int slot = gd_grad_slot(p);
    float g = 0.f;
    for (int b = 0; b < blocks; ++b) g += partial[(size_t)b * GD_MLP_PARTIAL + slot];
    grad[p] = g;
    if (doAdam) {
        const float m = beta1 * m1[p] + (1.f - beta1) * g;
        const float v = beta2 * m2[p] + (1.f - beta2) * g * g;
        m1[p] = m;
        m2[p] = v;
        const float mh = m / c1, vh = v / c2;
        master[p] -= lr * mh / (sqrtf(vh) + eps);
    }
}

/* ---------------------------------------------------------------------------
 * Host API
 * ------------------------------------------------------------------------- */
struct gd_mlp {
    float* master;       /* FP32 parameters */
    float* m1;
    float* m2;
    float* grad;
    gd_bf16x4* fwd;      /* 40 tiles x 64 lanes */
    gd_bf16x4* bwd;      /* 36 tiles x 64 lanes */
    float* bias;         /* GD_MLP_BIAS_PAD */
    float* partial;
    float* lossPart;
    int partialBlocks;
    int step;
};

static void gd_mlp_repack(gd_mlp* net) {
    k_mlp_pack<<<GD_MLP_FWD_TILES + GD_MLP_BWD_TILES, 64, 0, gd_stream()>>>(net->master, net->fwd, net->bwd,
                                                            net->bias);
    GD_LAUNCHED();
}

extern "C" {

gd_mlp* gd_mlp_create(const float* hostParams) {
    gd_mlp* net = (gd_mlp*)calloc(1, sizeof(gd_mlp));
    if (!net) {
        GD_FAIL("gd_mlp_create: out of host memory");
        return nullptr;
    }
    const size_t pb = sizeof(float) * GD_MLP_NPARAMS;
    GD_HIP(hipMalloc(&net->master, pb));
    GD_HIP(hipMalloc(&net->m1, pb));
    GD_HIP(hipMalloc(&net->m2, pb));
    GD_HIP(hipMalloc(&net->grad, pb));
    GD_HIP(hipMalloc(&net->fwd, sizeof(gd_bf16x4) * 64 * GD_MLP_FWD_TILES));
    GD_HIP(hipMalloc(&net->bwd, sizeof(gd_bf16x4) * 64 * GD_MLP_BWD_TILES));
    GD_HIP(hipMalloc(&net->bias, sizeof(float) * GD_MLP_BIAS_PAD));
    GD_HIP(hipMemset(net->m1, 0, pb));
    GD_HIP(hipMemset(net->m2, 0, pb));
    if (hostParams) GD_HIP(hipMemcpy(net->master, hostParams, pb, hipMemcpyHostToDevice));
    else GD_HIP(hipMemset(net->master, 0, pb));
    if (gd_has_error()) return net;
    gd_mlp_repack(net);
    return net;
}

void gd_mlp_destroy(gd_mlp* net) {
    if (!net) return;
    GD_HIP(hipFree(net->master));
    GD_HIP(hipFree(net->m1));
    GD_HIP(hipFree(net->m2));
    GD_HIP(hipFree(net->grad));
    GD_HIP(hipFree(net->fwd));
    GD_HIP(hipFree(net->bwd));
    GD_HIP(hipFree(net->bias));
    if (net->partial) GD_HIP(hipFree(net->partial));
    if (net->lossPart) GD_HIP(hipFree(net->lossPart));
    free(net);
}

void gd_mlp_get_params(const gd_mlp* net, float* hostParams) {
    GD_HIP(hipMemcpy(hostParams, net->master, sizeof(float) * GD_MLP_NPARAMS,
                     hipMemcpyDeviceToHost));
}

void gd_mlp_set_params(gd_mlp* net, const float* hostParams) {
    GD_HIP(hipMemcpy(net->master, hostParams, sizeof(float) * GD_MLP_NPARAMS,
                     hipMemcpyHostToDevice));
    gd_mlp_repack(net);
}

void gd_mlp_forward(const gd_mlp* net, const float* X, float* Y, long long rows) {
    if (rows <= 0) return;
    long long blocks = (rows + 15) / 16;
    if (blocks > 16384) blocks = 16384;
    k_mlp_infer<<<(int)blocks, 64, 0, gd_stream()>>>(net->fwd, net->bias, X, Y, rows);
    GD_LAUNCHED();
}

static float gd_mlp_grad_impl(gd_mlp* net, const float* X, const float* Y, const int* order,
                              long long B, float* hostGrad, int doAdam, float lr) {
    const int perBlock = 256;
    const int blocks = (int)((B + perBlock - 1) / perBlock);
    if (blocks > net->partialBlocks) {
        if (net->partial) GD_HIP(hipFree(net->partial));
        if (net->lossPart) GD_HIP(hipFree(net->lossPart));
        GD_HIP(hipMalloc(&net->partial, sizeof(float) * (size_t)GD_MLP_PARTIAL * blocks));
        GD_HIP(hipMalloc(&net->lossPart, sizeof(float) * blocks));
        net->partialBlocks = blocks;
    }
    if (gd_has_error()) return -1.f;
    k_mlp_grad<<<blocks, 64, 0, gd_stream()>>>(net->fwd, net->bwd, net->bias, X, Y, order, B, perBlock,
                               net->partial, net->lossPart);
    GD_LAUNCHED();
    float c1 = 1.f, c2 = 1.f;
    if (doAdam) {
        net->step += 1;
        c1 = 1.f - powf(0.9f, (float)net->step);
        c2 = 1.f - powf(0.999f, (float)net->step);
    }
    k_mlp_reduce_adam<<<gd_div_up(GD_MLP_NPARAMS, 256), 256, 0, gd_stream()>>>(
        net->partial, blocks, net->grad, net->master, net->m1, net->m2, doAdam, lr, 0.9f, 0.999f,
        1e-8f, c1, c2);
    GD_LAUNCHED();
    if (doAdam) gd_mlp_repack(net);
    if (hostGrad)
        GD_HIP(hipMemcpy(hostGrad, net->grad, sizeof(float) * GD_MLP_NPARAMS,
                         hipMemcpyDeviceToHost));
    float* lp = (float*)malloc(sizeof(float) * blocks);
    GD_HIP(hipMemcpy(lp, net->lossPart, sizeof(float) * blocks, hipMemcpyDeviceToHost));
    double loss = 0.0;
    for (int b = 0; b < blocks; ++b) loss += lp[b];
    free(lp);
    return (float)(loss / (double)B);
}

float gd_mlp_gradient(gd_mlp* net, const float* X, const float* Y, const int* order, long long B,
                      float* hostGrad) {
    return gd_mlp_grad_impl(net, X, Y, order, B, hostGrad, 0, 0.f);
}

float gd_mlp_train_step(gd_mlp* net, const float* X, const float* Y, const int* order,
                        long long B, float lr) {
    return gd_mlp_grad_impl(net, X, Y, order, B, nullptr, 1, lr);
}

void gd_mlp_adam_state(const gd_mlp* net, float* hostM, float* hostV) {
    if (hostM) GD_HIP(hipMemcpy(hostM, net->m1, sizeof(float) * GD_MLP_NPARAMS, hipMemcpyDeviceToHost));
    if (hostV) GD_HIP(hipMemcpy(hostV, net->m2, sizeof(float) * GD_MLP_NPARAMS, hipMemcpyDeviceToHost));
}

}  /* extern "C" */

/* internal accessors for the walk */
const gd_bf16x4* gd_mlp_fwd_tiles(const gd_mlp* net) { return net->fwd; }
const float* gd_mlp_bias(const gd_mlp* net) { return net->bias; }

/* ---------------------------------------------------------------------------
 * Training samples: (features, scaled exact correction) of (body, cell) pairs
 * ------------------------------------------------------------------------- */
#define GD_SL_WAVES 4
#define GD_SL_STACK 2048

/* One wavefront per target body: the traversal the learned walk makes for
 * the body's group (its 64-body run, or the half of it, split by `ratio` as
 * the learned walk splits it; its bounding box against each node, the same
 * acceptance test); every accepted internal cell whose size / distance from
 * the group's centre exceeds minRatio -- the cells the learned walk corrects
 * -- becomes a (target, cell) pair, and centre[t] receives the group's
 * centre, where the walk evaluates the network. */
__global__ __launch_bounds__(64 * GD_SL_WAVES) void k_samples_list(
    const uint64_t* __restrict__ keys, const float4* __restrict__ pos, int n,
    const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const gd_tree_info* __restrict__ info,
    const int* __restrict__ targets, int nt, float theta, float ratio, float minRatio,
    int2* __restrict__ pairs,
    long long cap, unsigned long long* __restrict__ count, int* __restrict__ dropped,
    float4* __restrict__ centre) {
    __shared__ int sStack[GD_SL_WAVES][GD_SL_STACK];
    const unsigned lane = gd_lane();
    const int wave = threadIdx.x / 64;
    const int t = blockIdx.x * GD_SL_WAVES + wave;
    if (t >= nt) return;
    int* stack = sStack[wave];
    const int b = targets[t];
    const long long run = b / 64;
    const int j = gd_split_run(keys, pos, n, run, ratio);
    const int rstart = (int)(run * 64), rcount = n - rstart < 64 ? n - rstart : 64;
    const bool second = j > 0 && b - rstart >= j;
    const int gstart = second ? rstart + j : rstart;
    const int gcount = j == 0 ? rcount : (second ? rcount - j : j);
    float lx, ly, lz, hx, hy, hz;
    gd_range_box(pos, gstart, gcount, lx, ly, lz, hx, hy, hz);
    const float gx = 0.5f * (lx + hx), gy = 0.5f * (ly + hy), gz = 0.5f * (lz + hz);
    if (lane == 0) centre[t] = make_float4(gx, gy, gz, 0.f);
    const float mr2 = minRatio > 0.f ? minRatio * minRatio : 0.f;
    const float ext = info->box.extent;
    const float th2 = theta * theta;
    const unsigned long long lt = gd_lanemask_lt();
    int top = 1;
    if (lane == 0) stack[0] = 0;
    while (top > 0) {
        const int take = top < 64 ? top : 64;
        const int base = top - take;
        gd_probe pr = {0, 0, 0, false, false, false};
        if ((int)lane < take) {
            const int idx = stack[base + lane];
            pr = gd_test_node(wcom[idx], wlink[idx], idx, ext, th2, lx, ly, lz, hx, hy, hz,
                              wrad != nullptr, wrad ? wrad[idx] : 0.f);
        }
        __builtin_amdgcn_wave_barrier();
        top = base;
        bool cell = pr.accept && !pr.leaf;
        if (cell) {
            /* the walk's correction test, with its cell measure (the mass
               radius under the mass-radius rule, else the cell size) */
            const float4 c = wcom[pr.idx];
            const float size = wrad ? wrad[pr.idx] : ldexpf(ext, -(wlink[pr.idx].y & 31));
            const float dx = c.x - gx, dy = c.y - gy, dz = c.z - gz;
            cell = size * size > mr2 * (dx * dx + dy * dy + dz * dz);
        }
        const unsigned long long bal = __ballot(cell);
        const int nacc = __popcll(bal);
        unsigned long long slot0 = 0;
        if (lane == 0 && nacc) slot0 = atomicAdd(count, (unsigned long long)nacc);
        slot0 = __shfl(slot0, 0);
        if (cell) {
            const unsigned long long slot = slot0 + (unsigned long long)__popcll(bal & lt);
            if ((long long)slot < cap) pairs[slot] = make_int2(t, pr.idx);
        }
        int excl, total;
        gd_wave_scan16(pr.nk, lt, excl, total);
        if (top + total > GD_SL_STACK) {
            if (lane == 0) atomicAdd(dropped, 1);
            break;
        }
        if (pr.open)
            for (int k = 0; k < pr.nk; ++k) stack[top + excl + k] = pr.first + k;
        top += total;
        __builtin_amdgcn_wave_barrier();
    }
}

/* One wavefront per pair: the 8 features (the walk's: u, s/r,
 * (7.5 O:uu - 1.5 T)/(m s^3), u.O:uu/(m s^3)) and the target (sum over the
 * cell's bodies - softened monopole - analytic quadrupole) / (m s^3 / r^5),
 * in FP64: what the quadrupole cell misses, which the network learns. */
__global__ __launch_bounds__(256) void k_samples_eval(
    const float4* __restrict__ pos, const int4* __restrict__ meta,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wcom, const float4* __restrict__ wmom,
    const gd_tree_info* __restrict__ info, const int* __restrict__ targets,
    const float4* __restrict__ centre, const int2* __restrict__ pairs, long long npairs,
    double eps2, float* __restrict__ X, float* __restrict__ Y) {
    const unsigned lane = gd_lane();
    const long long waves = (long long)gridDim.x * 4;
    for (long long p = (long long)blockIdx.x * 4 + threadIdx.x / 64; p < npairs; p += waves) {
        const int2 pr = pairs[p];
        const float4 x = pos[targets[pr.x]];
        const int idx = pr.y;
        const int4 mt = meta[idx];   /* first body, body count (octant in bits 29-31) */
        const int first = mt.x;
        const int count = mt.y & 0x1fffffff;
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j = lane; j < count; j += 64) {
            const float4 b = pos[first + j];
            const double dx = (double)b.x - x.x, dy = (double)b.y - x.y, dz = (double)b.z - x.z;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double f = (double)b.w / (r2 * sqrt(r2));
            ax += f * dx;
            ay += f * dy;
            az += f * dz;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            ax += __shfl_xor(ax, off);
            ay += __shfl_xor(ay, off);
            az += __shfl_xor(az, off);
        }
        if (lane == 0) {
            const float4 c = wcom[idx];
            float4 q0, q1;
            gd_mom_quad(wmom, idx, q0, q1);
            /* the cell measure s, as the walk takes it */
            const double size = wrad ? (double)wrad[idx]
                                     : ldexp((double)info->box.extent, -(wlink[idx].y & 31));
            const double dx = (double)c.x - x.x, dy = (double)c.y - x.y, dz = (double)c.z - x.z;
            const double rc2 = dx * dx + dy * dy + dz * dz;
            const double rs2 = rc2 + eps2;
            const double mono = (double)c.w / (rs2 * sqrt(rs2));
            const double scale = (double)c.w * size * size * size / (rc2 * rc2 * sqrt(rc2));
            /* analytic quadrupole term, as the walk evaluates it (softened r) */
            const double sxx = q0.x, syy = q0.y, szz = q0.z, sxy = q0.w, sxz = q1.x, syz = q1.y;
            const double tr = q1.z;
            const double sdx = sxx * dx + sxy * dy + sxz * dz;
            const double sdy = sxy * dx + syy * dy + syz * dz;
            const double sdz = sxz * dx + syz * dy + szz * dz;
            const double q = dx * sdx + dy * sdy + dz * sdz;
            const double inv5 = 1.0 / (rs2 * rs2 * sqrt(rs2));
            const double cd = inv5 * (7.5 * q / rs2 - 1.5 * tr);
            const double qx = cd * dx - 3.0 * inv5 * sdx;
            const double qy = cd * dy - 3.0 * inv5 * sdy;
            const double qz = cd * dz - 3.0 * inv5 * sdz;
            Y[p * 3 + 0] = (float)((ax - mono * dx - qx) / scale);
            Y[p * 3 + 1] = (float)((ay - mono * dy - qy) / scale);
            Y[p * 3 + 2] = (float)((az - mono * dz - qz) / scale);
            /* features from the group's centre, as the walk computes them */
            const float4 g = centre[pr.x];
            const double gx = (double)c.x - g.x, gy = (double)c.y - g.y, gz = (double)c.z - g.z;
            const double r = sqrt(gx * gx + gy * gy + gz * gz);
            const double ux = gx / r, uy = gy / r, uz = gz / r;
            const double k3 = (double)c.w * size * size * size;
            float4 o0, o1, o2;
            gd_mom_oct(wmom, idx, o0, o1, o2);
            const double xx = ux * ux, yy = uy * uy, zz = uz * uz;
            const double xy2 = 2.0 * ux * uy, xz2 = 2.0 * ux * uz, yz2 = 2.0 * uy * uz;
            /* O_ijk u_j u_k and the trace vector T_i = O_ijj */
            const double ox = o0.x * xx + o1.y * yy + o1.w * zz + o0.w * xy2 + o1.x * xz2 +
                              o2.y * yz2;
            const double oy = o0.w * xx + o0.y * yy + o2.x * zz + o1.y * xy2 + o2.y * xz2 +
                              o1.z * yz2;
            const double oz = o1.x * xx + o1.z * yy + o0.z * zz + o2.y * xy2 + o1.w * xz2 +
                              o2.x * yz2;
            const double tx = (double)o0.x + o1.y + o1.w;
            const double ty = (double)o0.w + o0.y + o2.x;
            const double tz = (double)o1.x + o1.z + o0.z;
            X[p * 8 + 0] = (float)ux;
            X[p * 8 + 1] = (float)uy;
            X[p * 8 + 2] = (float)uz;
            X[p * 8 + 3] = (float)(size / r);
            X[p * 8 + 4] = (float)((7.5 * ox - 1.5 * tx) / k3);
            X[p * 8 + 5] = (float)((7.5 * oy - 1.5 * ty) / k3);
            X[p * 8 + 6] = (float)((7.5 * oz - 1.5 * tz) / k3);
            X[p * 8 + 7] = (float)((ux * ox + uy * oy + uz * oz) / k3);
        }
    }
}

extern "C" long long gd_corrector_samples(const treeState* tree, const float4* pos,
                                          const int* targets, int nt, float theta,
                                          float minRatio, float eps2, float* X, float* Y,
                                          long long cap, int* droppedTargets) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_corrector_samples: uninitialised tree state");
        return 0;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    if (!(b->quad == 2 && b->wmom)) {
        GD_FAIL("gd_corrector_samples: build the tree with second and third moments first "
                "(gd_set_quadrupole(tree, 2))");
        return 0;
    }
    if (nt <= 0 || cap <= 0) return 0;
    {
        int overflow = 0;
        GD_HIP(hipMemcpy(&overflow, &b->dinfo->overflow, sizeof(int), hipMemcpyDeviceToHost));
        if (overflow) {   /* basicGravity.cu:282 stops on the same condition */
            GD_FAIL("Node array overflow: the tree does not fit its %d nodes", b->maxNodes);
            return 0;
        }
    }
    int2* pairs = nullptr;
    unsigned long long* count = nullptr;
    int* dropped = nullptr;
    float4* centre = nullptr;
    GD_HIP(hipMalloc(&pairs, sizeof(int2) * (size_t)cap));
    GD_HIP(hipMalloc(&count, sizeof(unsigned long long)));
    GD_HIP(hipMalloc(&dropped, sizeof(int)));
    GD_HIP(hipMalloc(&centre, sizeof(float4) * (size_t)nt));
    GD_HIP(hipMemset(count, 0, sizeof(unsigned long long)));
    GD_HIP(hipMemset(dropped, 0, sizeof(int)));
    long long n = 0;
    if (!gd_has_error()) {
        k_samples_list<<<gd_div_up(nt, GD_SL_WAVES), 64 * GD_SL_WAVES, 0, gd_stream()>>>(
            tree->mortonOut, pos, tree->starCount, b->wcom, b->wlink, gd_walk_radii(b), b->dinfo,
            targets, nt, theta,
            gd_walk_split_ratio(theta), minRatio, pairs, cap, count, dropped, centre);
        GD_LAUNCHED();
        unsigned long long total = 0;
        int drop = 0;
        GD_HIP(hipMemcpy(&total, count, sizeof(total), hipMemcpyDeviceToHost));
        GD_HIP(hipMemcpy(&drop, dropped, sizeof(int), hipMemcpyDeviceToHost));
        const long long rows = (long long)total < cap ? (long long)total : cap;
        n = (long long)total;
        if (droppedTargets) *droppedTargets = drop;
        if (rows > 0) {
            long long blocks = (rows + 3) / 4;
            if (blocks > 65536) blocks = 65536;
            k_samples_eval<<<(int)blocks, 256, 0, gd_stream()>>>(pos, b->meta, b->wlink,
                                                 gd_walk_radii(b), b->wcom, b->wmom,
                                                 b->dinfo, targets, centre, pairs, rows,
                                                 (double)eps2, X, Y);
            GD_LAUNCHED();
        }
        GD_HIP(hipDeviceSynchronize());
    }
    GD_HIP(hipFree(centre));
    GD_HIP(hipFree(pairs));
    GD_HIP(hipFree(count));
    GD_HIP(hipFree(dropped));
    return n;
}
