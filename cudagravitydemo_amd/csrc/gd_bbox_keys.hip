/*
 * gd_bbox_keys.hip -- bounding box, quantisation box, Morton keys, reorder.
 *
 * Reference behaviour restated:
 *   bbox          tree.cu:385-490   exact FP32 per-axis min/max
 *   world box     basicGravity.cu:90-112   extent = max width, scale = (2^21-1)/extent
 *   keys          tree.cu:332-367   q = (u32)((x - min) * scale), x at bit 3k+2,
 *                                    y at 3k+1, z at 3k; payload = iota
 *   reorder       sorting.cu:99-107 gather by the sorted payload
 *
 * Design: the engine's bbox is a two-kernel reduction (GD_BBOX_BLOCKS
 * partials with wave-level shuffles, then one block) that also computes the
 * quantisation box on the device, so a step never waits for the host.  The
 * key kernel reads that box from device memory.  All kernels stream with one
 * body per thread and coalesced 16-byte position reads.
 */
#include "gd_internal.h"

#include <math.h>

/* ---------------------------------------------------------------------------
 * Bounding box
 * ------------------------------------------------------------------------- */
struct gd_minmax {
    float lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ void gd_minmax_wave(gd_minmax& m) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m.lx = fminf(m.lx, __shfl_xor(m.lx, off));
        m.ly = fminf(m.ly, __shfl_xor(m.ly, off));
        m.lz = fminf(m.lz, __shfl_xor(m.lz, off));
        m.hx = fmaxf(m.hx, __shfl_xor(m.hx, off));
        m.hy = fmaxf(m.hy, __shfl_xor(m.hy, off));
        m.hz = fmaxf(m.hz, __shfl_xor(m.hz, off));
    }
}

/* Block-wide min/max of m; the result is valid in thread 0. */
template <int THREADS>
__device__ __forceinline__ void gd_minmax_block(gd_minmax& m) {
    __shared__ gd_minmax part[THREADS / GD_WAVE];
    gd_minmax_wave(m);
    const int wave = threadIdx.x / GD_WAVE;
    if (gd_lane() == 0) part[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < THREADS / GD_WAVE; ++w) {
            m.lx = fminf(m.lx, part[w].lx); m.ly = fminf(m.ly, part[w].ly);
            m.lz = fminf(m.lz, part[w].lz); m.hx = fmaxf(m.hx, part[w].hx);
            m.hy = fmaxf(m.hy, part[w].hy); m.hz = fmaxf(m.hz, part[w].hz);
        }
    }
}

__device__ __forceinline__ gd_minmax gd_minmax_empty() {
    gd_minmax m;
    m.lx = m.ly = m.lz = INFINITY;
    m.hx = m.hy = m.hz = -INFINITY;
    return m;
}

/* FP64 mass and mass-weighted position of one body, hi + lo parts. */
__device__ __forceinline__ void gd_com_add(double* s, float4 p, float3 l) {
    const double x = (double)p.x + l.x, y = (double)p.y + l.y, z = (double)p.z + l.z;
    const double m = p.w;
    s[0] += m;
    s[1] = __builtin_fma(m, x, s[1]);
    s[2] = __builtin_fma(m, y, s[2]);
    s[3] = __builtin_fma(m, z, s[3]);
}

/* Per-block partial min/max over a grid-stride range of bodies and, with
 * COM, the block's FP64 sums m, m x, m y, m z (pcom[4 block ..]) over hi + lo
 * positions (LO: lo is read, else the low parts are zero).  Every load of an
 * iteration is issued before any is used. */
template <bool COM, bool LO>
__global__ __launch_bounds__(GD_BBOX_THREADS) void k_bbox_partial(const float4* __restrict__ pos,
                                                                  const float3* __restrict__ lo,
                                                                  int n, float3* pmin,
                                                                  float3* pmax, double* pcom) {
    __shared__ double sCom[GD_BBOX_THREADS / 64][4];
    gd_minmax m = gd_minmax_empty();
    double cs[4] = {0.0, 0.0, 0.0, 0.0};
    const float3 z3 = make_float3(0.f, 0.f, 0.f);
    const int stride = gridDim.x * GD_BBOX_THREADS;
    int i = blockIdx.x * GD_BBOX_THREADS + threadIdx.x;
    /* two bodies per iteration keep two 16-byte loads in flight */
    for (; i + stride < n; i += 2 * stride) {
        const float4 a = pos[i];
        const float4 b = pos[i + stride];
        const float3 la = LO ? lo[i] : z3;
        const float3 lb = LO ? lo[i + stride] : z3;
        m.lx = fminf(m.lx, fminf(a.x, b.x)); m.hx = fmaxf(m.hx, fmaxf(a.x, b.x));
        m.ly = fminf(m.ly, fminf(a.y, b.y)); m.hy = fmaxf(m.hy, fmaxf(a.y, b.y));
        m.lz = fminf(m.lz, fminf(a.z, b.z)); m.hz = fmaxf(m.hz, fmaxf(a.z, b.z));
        if (COM) {
            gd_com_add(cs, a, la);
            gd_com_add(cs, b, lb);
        }
    }
    if (i < n) {
        const float4 a = pos[i];
        const float3 la = LO ? lo[i] : z3;
        m.lx = fminf(m.lx, a.x); m.hx = fmaxf(m.hx, a.x);
        m.ly = fminf(m.ly, a.y); m.hy = fmaxf(m.hy, a.y);
        m.lz = fminf(m.lz, a.z); m.hz = fmaxf(m.hz, a.z);
        if (COM) gd_com_add(cs, a, la);
    }
    gd_minmax_block<GD_BBOX_THREADS>(m);
    if (COM) {   /* block sum: wavefront butterflies, then the wavefronts in order */
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) cs[k] += __shfl_xor(cs[k], off);
        }
        if (gd_lane() == 0)
            for (int k = 0; k < 4; ++k) sCom[threadIdx.x >> 6][k] = cs[k];
        __syncthreads();
        if (threadIdx.x < 4) {
            double t = 0.0;
            for (int w = 0; w < GD_BBOX_THREADS / 64; ++w) t += sCom[w][threadIdx.x];
            pcom[4 * blockIdx.x + threadIdx.x] = t;
        }
    }
    if (threadIdx.x == 0) {
        pmin[blockIdx.x] = make_float3(m.lx, m.ly, m.lz);
        pmax[blockIdx.x] = make_float3(m.hx, m.hy, m.hz);
    }
}

/*
 * Reduce `count` partials into outMin/outMax; with info != NULL also store
 * the corners and the quantisation box there (basicGravity.cu:90-112 in
 * device float arithmetic: the same two IEEE operations the host did).
 */
__global__ __launch_bounds__(1024) void k_bbox_final(const float3* __restrict__ pmin,
                                                     const float3* __restrict__ pmax, int count,
                                                     float3* outMin, float3* outMax,
                                                     gd_tree_info* info,
                                                     const double* __restrict__ pcom) {
    __shared__ double sSum[1024][4];
    if (pcom && info) {   /* uniform: the blocks' FP64 sums by a fixed pairwise tree */
        const int t = threadIdx.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) sSum[t][k] = t < count ? pcom[4 * t + k] : 0.0;
        __syncthreads();
        for (int h = 512; h > 0; h >>= 1) {
            if (t < h)
#pragma unroll
                for (int k = 0; k < 4; ++k) sSum[t][k] += sSum[t + h][k];
            __syncthreads();
        }
        if (t == 0) {
            const double mt = sSum[0][0];
            info->rootCom[0] = mt > 0.0 ? sSum[0][1] / mt : 0.0;
            info->rootCom[1] = mt > 0.0 ? sSum[0][2] / mt : 0.0;
            info->rootCom[2] = mt > 0.0 ? sSum[0][3] / mt : 0.0;
            info->rootCom[3] = mt;
        }
    }
    gd_minmax m = gd_minmax_empty();
    for (int i = threadIdx.x; i < count; i += 1024) {
        const float3 a = pmin[i], b = pmax[i];
        m.lx = fminf(m.lx, a.x); m.ly = fminf(m.ly, a.y); m.lz = fminf(m.lz, a.z);
        m.hx = fmaxf(m.hx, b.x); m.hy = fmaxf(m.hy, b.y); m.hz = fmaxf(m.hz, b.z);
    }
    gd_minmax_block<1024>(m);
    if (threadIdx.x == 0) {
        const float3 lo = make_float3(m.lx, m.ly, m.lz);
        const float3 hi = make_float3(m.hx, m.hy, m.hz);
        *outMin = lo;
        *outMax = hi;
        if (info) {
            const float w = hi.x - lo.x, h = hi.y - lo.y, d = hi.z - lo.z;
            const float extent = fmaxf(w, fmaxf(h, d));
            info->box.minX = lo.x;
            info->box.minY = lo.y;
            info->box.minZ = lo.z;
            info->box.extent = extent;
            info->box.scale = (float)((1u << 21) - 1u) / extent;
            info->bmin = lo;
            info->bmax = hi;
        }
    }
}

void gd_launch_bbox(const float4* pos, const float3* lo, int n, gd_builder* b, float3* outMin,
                    float3* outMax) {
    const int blocks = n > 0 ? (gd_div_up(n, 2 * GD_BBOX_THREADS) < GD_BBOX_BLOCKS
                                    ? gd_div_up(n, 2 * GD_BBOX_THREADS)
                                    : GD_BBOX_BLOCKS)
                             : 1;
    if (lo)
        k_bbox_partial<true, true><<<blocks, GD_BBOX_THREADS, 0, gd_stream()>>>(
            pos, lo, n, b->bboxPartMin, b->bboxPartMax, b->bboxPartCom);
    else
        k_bbox_partial<true, false><<<blocks, GD_BBOX_THREADS, 0, gd_stream()>>>(
            pos, nullptr, n, b->bboxPartMin, b->bboxPartMax, b->bboxPartCom);
    GD_LAUNCHED();
    k_bbox_final<<<1, 1024, 0, gd_stream()>>>(b->bboxPartMin, b->bboxPartMax, blocks, outMin,
                                              outMax, b->dinfo, b->bboxPartCom);
    GD_LAUNCHED();
}

/* ---------------------------------------------------------------------------
 * Morton keys
 * ------------------------------------------------------------------------- */

/* Bits 0..6 of v moved to bits 0,3,6,...,18. */
__device__ __forceinline__ uint32_t gd_spread7(uint32_t v) {
    v &= 0x7fu;
    v = (v | (v << 8)) & 0x0000f00fu;   /* b0-b3 stay, b4-b6 -> 12-14 */
    v = (v | (v << 4)) & 0x000c30c3u;   /* pairs at 0,6,12,18 */
    v = (v | (v << 2)) & 0x00249249u;   /* single bits at multiples of 3 */
    return v;
}

/* Bits 0..20 of v moved to bits 0,3,...,60 (three 7-bit chunks). */
__device__ __forceinline__ uint64_t gd_spread21(uint32_t v) {
    return (uint64_t)gd_spread7(v) | ((uint64_t)gd_spread7(v >> 7) << 21) |
           ((uint64_t)gd_spread7(v >> 14) << 42);
}

/* The reference's (uint32_t) cast as its GPU executes it: round toward zero,
 * saturate, NaN -> 0 (a zero-extent box gives NaN coordinates). */
__device__ __forceinline__ uint32_t gd_f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

__device__ __forceinline__ uint64_t gd_morton_key(float4 p, const worldBox& box) {
    const uint32_t qx = gd_f2u_sat((p.x - box.minX) * box.scale);
    const uint32_t qy = gd_f2u_sat((p.y - box.minY) * box.scale);
    const uint32_t qz = gd_f2u_sat((p.z - box.minZ) * box.scale);
    return (gd_spread21(qx) << 2) | (gd_spread21(qy) << 1) | gd_spread21(qz);
}

__global__ __launch_bounds__(256) void k_morton_dev(const float4* __restrict__ pos, int n,
                                                    const gd_tree_info* __restrict__ info,
                                                    uint64_t* keys, uint32_t* idx) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const worldBox box = info->box;
    keys[i] = gd_morton_key(pos[i], box);
    idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_morton_val(const float4* __restrict__ pos, int n,
                                                    worldBox box, uint64_t* keys, uint32_t* idx) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    keys[i] = gd_morton_key(pos[i], box);
    idx[i] = (uint32_t)i;
}

void gd_launch_morton_dev(const float4* pos, int n, const gd_tree_info* info, uint64_t* keys,
                          uint32_t* idx) {
    if (n <= 0) return;
    k_morton_dev<<<gd_div_up(n, 256), 256, 0, gd_stream()>>>(pos, n, info, keys, idx);
    GD_LAUNCHED();
}

/* ---------------------------------------------------------------------------
 * Reorder
 * ------------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_reorder(const float4* __restrict__ posIn,
                                                 const float3* __restrict__ velIn,
                                                 float4* __restrict__ posOut,
                                                 float3* __restrict__ velOut,
                                                 const uint32_t* __restrict__ idx, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = idx[i];
    posOut[i] = posIn[j];
    velOut[i] = velIn[j];
}

__global__ __launch_bounds__(256) void k_reorder_lo(const float4* __restrict__ posIn,
                                                    const float3* __restrict__ velIn,
                                                    const float3* __restrict__ loIn,
                                                    float4* __restrict__ posOut,
                                                    float3* __restrict__ velOut,
                                                    float3* __restrict__ loOut,
                                                    const uint32_t* __restrict__ idx, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = idx[i];
    posOut[i] = posIn[j];
    velOut[i] = velIn[j];
    loOut[i] = loIn[j];
}

void gd_launch_reorder(const float4* posIn, const float3* velIn, float4* posOut, float3* velOut,
                       const uint32_t* idx, int n) {
    if (n <= 0) return;
    k_reorder<<<gd_div_up(n, 256), 256, 0, gd_stream()>>>(posIn, velIn, posOut, velOut, idx, n);
    GD_LAUNCHED();
}

void gd_launch_reorder_lo(const float4* posIn, const float3* velIn, const float3* loIn,
                          float4* posOut, float3* velOut, float3* loOut, const uint32_t* idx,
                          int n) {
    if (n <= 0) return;
    k_reorder_lo<<<gd_div_up(n, 256), 256, 0, gd_stream()>>>(posIn, velIn, loIn, posOut, velOut, loOut, idx, n);
    GD_LAUNCHED();
}

/* ---------------------------------------------------------------------------
 * Reference-API launchers
 * ------------------------------------------------------------------------- */
extern "C" {

void mortonEncode(const float4* posMassVals, uint64_t* mortonCodes, uint32_t* originalIndex,
                  int NStars, worldBox rootBox) {
    if (NStars <= 0) return;
    k_morton_val<<<gd_div_up(NStars, 256), 256, 0, gd_stream()>>>(posMassVals, NStars, rootBox, mortonCodes,
                                                  originalIndex);
    GD_LAUNCHED();
}

void localMinMaxFinder(const float4* posMassVals, unsigned int N, float3* minCorner,
                       float3* maxCorner) {
    /* BOXBLOCKS x BOXTHREADS like tree.h:16-17; each block writes one partial */
    k_bbox_partial<false, false><<<256, GD_BBOX_THREADS, 0, gd_stream()>>>(posMassVals, nullptr, (int)N, minCorner,
                                                               maxCorner, nullptr);
    GD_LAUNCHED();
}

void globalMinMaxReducer(const float3* minCorner, const float3* maxCorner, int count,
                         float3* outMin, float3* outMax) {
    k_bbox_final<<<1, 1024, 0, gd_stream()>>>(minCorner, maxCorner, count, outMin, outMax, nullptr,
                                              nullptr);
    GD_LAUNCHED();
}

void reorderParticles(const float4* posMassValsIn, const float3* velValsIn,
                      float4* posMassValsOut, float3* velValsOut, uint32_t* originalIndex,
                      int starCount) {
    gd_launch_reorder(posMassValsIn, velValsIn, posMassValsOut, velValsOut, originalIndex,
                      starCount);
}

}  /* extern "C" */
