/*
 * gd_walk.hip -- tree walks.
 *
 * 1. Parity walk (calculateAcceleration, tree.cu:562-676).  The reference
 *    walks one 32-body group per 32-thread warp: a LIFO stack shared by the
 *    group, a node is opened when any valid body fails size^2 < theta^2 r^2 on
 *    a non-leaf (size = bbox x extent / 2^level), otherwise every body adds
 *    the node's softened monopole.  A 64-wide wavefront here carries two such
 *    groups, one per half, each with its own LDS stack and its own 32-bit
 *    slice of the wavefront ballot, so each group's interaction list is the
 *    reference's.  The arithmetic follows the oracle's operation order
 *    (explicit fmaf, IEEE division and square root, no contraction).
 *
 * 2. Production walk (gd_walk).  One 64-body group per wavefront (consecutive
 *    bodies in Morton order).  The group's bounding box is tested against up
 *    to 64 nodes at once, one node per lane: a node is accepted when it is a
 *    leaf or size^2 < theta^2 d^2, d being the distance from the group box to
 *    the node's centre of mass and size = bbox cube extent / 2^level; opened
 *    nodes push their (contiguous) children on an LDS stack.  Accepted nodes
 *    go to an LDS interaction list; when it fills, all 64 lanes evaluate it
 *    with broadcast LDS reads (compute bound: ~20 FP32 FLOP per entry).  Per
 *    body, each flush's partial sum is added to the total with Kahan
 *    compensation.  A group whose stack would overflow its LDS share is
 *    re-walked by the same code with its stack in global memory.
 */
#include "gd_internal.h"
#include "gd_mlp.h"

#include <math.h>

/* ---------------------------------------------------------------------------
 * Parity walk
 * ------------------------------------------------------------------------- */
#define GD_PAR_STACK 168          /* 8 x 21, as tree.cu:588 */
#define GD_PAR_SLOTS 8            /* 32-body groups per 256-thread block */

__global__ __launch_bounds__(256) void k_walk_parity(const float4* __restrict__ pos,
                                                     const node* __restrict__ nodes,
                                                     const float3* __restrict__ bmin,
                                                     const float3* __restrict__ bmax,
                                                     float3* __restrict__ acc, float eps2,
                                                     float G, float theta, int n) {
    __shared__ int sStack[GD_PAR_SLOTS][GD_PAR_STACK];
    const unsigned lane = gd_lane();
    const int half = (int)(lane >> 5);
    const int slot = (int)(threadIdx.x >> 6) * 2 + half;
    int* stack = sStack[slot];
    const long long g = (long long)blockIdx.x * GD_PAR_SLOTS + slot;
    const long long i = g * 32 + (lane & 31u);
    const bool valid = i < n;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (valid) {
        const float4 p = pos[i];
        px = p.x; py = p.y; pz = p.z;
    }
    const float ext = bmax->x - bmin->x;
    const float th2 = theta * theta;
    float ax = 0.f, ay = 0.f, az = 0.f;

    int top = -1;
    {
        const int4* r = reinterpret_cast<const int4*>(nodes);
        const int4 w1 = r[1], w2 = r[2], w3 = r[3];
        const int ch[8] = {w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w, w3.x};
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (ch[c] != -1) stack[++top] = ch[c];   /* every lane of the half writes the same value */
    }
    if (g * 32 >= n) top = -1;   /* slot past the last group */

    while (__any(top >= 0)) {
        const bool active = top >= 0;
        bool want = false;
        float dx = 0.f, dy = 0.f, dz = 0.f, d2 = 0.f, mass = 0.f;
        int ch[8];
        if (active) {
            const int idx = stack[top];
            const int4* r = reinterpret_cast<const int4*>(nodes + idx);
            const int4 w1 = r[1], w2 = r[2], w3 = r[3];
            const float4 cm = nodes[idx].massData;
            ch[0] = w1.y; ch[1] = w1.z; ch[2] = w1.w; ch[3] = w2.x;
            ch[4] = w2.y; ch[5] = w2.z; ch[6] = w2.w; ch[7] = w3.x;
            const int count = w1.x, level = w3.y;
            dx = px - cm.x;
            dy = py - cm.y;
            dz = pz - cm.z;
            d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            mass = cm.w;
            const float size = ldexpf(ext, -level);   /* exact: ext / 2^level */
            int kids = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) kids += ch[c] != -1;
            const bool leaf = (count <= 1) || (kids == 0);
            const bool accept = (size * size) < th2 * d2;
            want = !(accept || !valid || leaf);
        }
        const unsigned long long bal = __ballot(want);
        const unsigned mine = (unsigned)(bal >> (32 * half));
        if (active) {
            --top;
            if (mine) {
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (ch[c] != -1) stack[++top] = ch[c];
            } else {
                const float r2 = d2 + eps2;
                const float mag = -(G * mass) / r2;
                const float dist = sqrtf(r2);
                ax = __builtin_fmaf(mag, dx / dist, ax);
                ay = __builtin_fmaf(mag, dy / dist, ay);
                az = __builtin_fmaf(mag, dz / dist, az);
            }
        }
    }
    if (valid) acc[i] = make_float3(ax, ay, az);
}

extern "C" void calculateAcceleration(const float4* positionMassArray, treeState tree,
                                      float3* accelVals, float antiSingularitySquared,
                                      float GVal, float theta, int starCount) {
    if (starCount <= 0) return;
    const int groups = gd_div_up(starCount, 32);
    k_walk_parity<<<gd_div_up(groups, GD_PAR_SLOTS), 256>>>(
        positionMassArray, tree.nodes, tree.boundingData[2], tree.boundingData[3], accelVals,
        antiSingularitySquared, GVal, theta, starCount);
    GD_LAUNCHED();
}

/* ---------------------------------------------------------------------------
 * Production walk
 * ------------------------------------------------------------------------- */
#ifndef GD_FW_WAVES
#define GD_FW_WAVES 4             /* wavefronts (64-body groups) per block */
#endif
#ifndef GD_FW_STACK
#define GD_FW_STACK 2048          /* LDS stack entries per wavefront */
#endif
#define GD_FW_BIG_STACK 12288     /* fallback LDS stack (48 KB): > 22 levels x 512 pushes */
#ifndef GD_FW_LIST
#define GD_FW_LIST 128            /* interaction-list entries per wavefront */
#endif

struct gd_kahan3 {
    float x, y, z, cx, cy, cz;
};

__device__ __forceinline__ void gd_kahan_add(float& s, float& c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

/* Monopole list[0..cnt) for one body; returns the partial sums. */
__device__ __forceinline__ void gd_eval_list(const float4* __restrict__ list, int cnt, float px,
                                             float py, float pz, float eps2, float& sx, float& sy,
                                             float& sz, float& sp) {
    float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
    int e = 0;
    for (; e + 2 <= cnt; e += 2) {
        const float4 c0 = list[e];
        const float4 c1 = list[e + 1];
        const float dx0 = c0.x - px, dy0 = c0.y - py, dz0 = c0.z - pz;
        const float dx1 = c1.x - px, dy1 = c1.y - py, dz1 = c1.z - pz;
        const float r0 = __builtin_fmaf(dx0, dx0, __builtin_fmaf(dy0, dy0, __builtin_fmaf(dz0, dz0, eps2)));
        const float r1 = __builtin_fmaf(dx1, dx1, __builtin_fmaf(dy1, dy1, __builtin_fmaf(dz1, dz1, eps2)));
        const float i0 = __builtin_amdgcn_rsqf(r0);
        const float i1 = __builtin_amdgcn_rsqf(r1);
        const float m0 = c0.w * i0, m1 = c1.w * i1;
        const float f0 = m0 * (i0 * i0), f1 = m1 * (i1 * i1);
        ax = __builtin_fmaf(f0, dx0, ax); ay = __builtin_fmaf(f0, dy0, ay); az = __builtin_fmaf(f0, dz0, az);
        ax = __builtin_fmaf(f1, dx1, ax); ay = __builtin_fmaf(f1, dy1, ay); az = __builtin_fmaf(f1, dz1, az);
        ph += m0;
        ph += m1;
    }
    if (e < cnt) {
        const float4 c0 = list[e];
        const float dx0 = c0.x - px, dy0 = c0.y - py, dz0 = c0.z - pz;
        const float r0 = __builtin_fmaf(dx0, dx0, __builtin_fmaf(dy0, dy0, __builtin_fmaf(dz0, dz0, eps2)));
        const float i0 = __builtin_amdgcn_rsqf(r0);
        const float m0 = c0.w * i0;
        const float f0 = m0 * (i0 * i0);
        ax = __builtin_fmaf(f0, dx0, ax); ay = __builtin_fmaf(f0, dy0, ay); az = __builtin_fmaf(f0, dz0, az);
        ph += m0;
    }
    sx = ax; sy = ay; sz = az; sp = ph;
}

/*
 * Quadrupole cell list for one body.  With d = c - x (toward the cell),
 * r^2 = |d|^2 + eps^2 and S the second moments about c:
 *   a/G = m d / r^3 + [ (15/2) (d.S d) / r^2 - (3/2) tr S ] d / r^5 - 3 S d / r^5
 *   phi/(-G) = m / r + (3 d.S d - r^2 tr S) / (2 r^5)
 */
__device__ __forceinline__ void gd_eval_quad(const float4* __restrict__ cm,
                                             const float4* __restrict__ q0,
                                             const float4* __restrict__ q1, int cnt, float px,
                                             float py, float pz, float eps2, float& sx, float& sy,
                                             float& sz, float& sp) {
    float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
    for (int e = 0; e < cnt; ++e) {
        const float4 c = cm[e];
        const float4 a = q0[e];   /* Sxx Syy Szz Sxy */
        const float4 b = q1[e];   /* Sxz Syz tr - */
        const float dx = c.x - px, dy = c.y - py, dz = c.z - pz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv2 = inv * inv;
        const float mi = c.w * inv;
        const float inv5 = inv * inv2 * inv2;
        const float sdx = __builtin_fmaf(a.x, dx, __builtin_fmaf(a.w, dy, b.x * dz));
        const float sdy = __builtin_fmaf(a.w, dx, __builtin_fmaf(a.y, dy, b.y * dz));
        const float sdz = __builtin_fmaf(b.x, dx, __builtin_fmaf(b.y, dy, a.z * dz));
        const float q = __builtin_fmaf(dx, sdx, __builtin_fmaf(dy, sdy, dz * sdz));
        const float cd = __builtin_fmaf(mi, inv2, inv5 * __builtin_fmaf(7.5f * q, inv2, -1.5f * b.z));
        const float cs = -3.f * inv5;
        ax = __builtin_fmaf(cd, dx, __builtin_fmaf(cs, sdx, ax));
        ay = __builtin_fmaf(cd, dy, __builtin_fmaf(cs, sdy, ay));
        az = __builtin_fmaf(cd, dz, __builtin_fmaf(cs, sdz, az));
        ph += __builtin_fmaf(0.5f * inv5, __builtin_fmaf(3.f, q, -r2 * b.z), mi);
    }
    sx = ax; sy = ay; sz = az; sp = ph;
}

#define GD_FW_QLIST 64            /* quadrupole / learned cell-list entries per wavefront */

/*
 * Learned correction of the cells in a list, evaluated once per group at the
 * group's centre g: for each cell, 8 features (unit vector to the COM, s/r,
 * S u / (m s^2), tr S / (m s^2)) go through the network on the matrix cores
 * (16 cells per pass, one cell per column), and the output, scaled by
 * m s^2 / r^4, is the correction of the cell's monopole.  Cells with
 * s/r <= minRatio are left uncorrected.  Returns the summed correction (without
 * the factor G) in every lane.
 */
__device__ __forceinline__ void gd_mlp_correct(const float4* __restrict__ qc,
                                               const float4* __restrict__ q0,
                                               const float4* __restrict__ q1, int cnt, float gx,
                                               float gy, float gz,
                                               const gd_bf16x4* __restrict__ W,
                                               const float* __restrict__ Bv, float minRatio,
                                               float& cx, float& cy, float& cz) {
    const unsigned l = gd_lane();
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int c0 = 0; c0 < cnt; c0 += 16) {
        const int cell = c0 + (int)(l % 16);
        bool ok = cell < cnt;
        float4 c = make_float4(gx + 1.f, gy, gz, 1.f);
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 b = make_float4(0.f, 0.f, 0.f, 1.f);
        if (ok) {
            const float4 cc = qc[cell];
            ok = cc.w > 0.f;   /* a massless cell needs no correction */
            if (ok) {
                c = cc;
                a = q0[cell];
                b = q1[cell];
            }
        }
        const float dx = c.x - gx, dy = c.y - gy, dz = c.z - gz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const float rinv = __builtin_amdgcn_rsqf(r2);
        const float ux = dx * rinv, uy = dy * rinv, uz = dz * rinv;
        const float size = b.w;
        const float ratio = size * rinv;
        const float k = c.w * size * size;
        const float kinv = 1.f / k;
        float f0, f1, f2, f3;
        const unsigned sel = l / 16;
        if (sel == 0) {
            f0 = ux; f1 = uy; f2 = uz; f3 = ratio;
        } else if (sel == 1) {
            f0 = __builtin_fmaf(a.x, ux, __builtin_fmaf(a.w, uy, b.x * uz)) * kinv;
            f1 = __builtin_fmaf(a.w, ux, __builtin_fmaf(a.y, uy, b.y * uz)) * kinv;
            f2 = __builtin_fmaf(b.x, ux, __builtin_fmaf(b.y, uy, a.z * uz)) * kinv;
            f3 = b.z * kinv;
        } else {
            f0 = f1 = f2 = f3 = 0.f;
        }
        const gd_f32x4 o = gd_mlp_infer16(W, Bv, gd_pack4(f0, f1, f2, f3));
        if (l < 16 && ok && ratio > minRatio) {
            const float r4inv = (rinv * rinv) * (rinv * rinv);
            const float scale = k * r4inv;
            sx = __builtin_fmaf(scale, o[0], sx);
            sy = __builtin_fmaf(scale, o[1], sy);
            sz = __builtin_fmaf(scale, o[2], sz);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
        sz += __shfl_xor(sz, off);
    }
    cx = sx; cy = sy; cz = sz;
}

/* MODE: 0 monopole, 1 quadrupole cells, 2 monopole cells + learned correction */
/* BIG: one wavefront per block walks a listed group with a GD_FW_BIG_STACK
 * LDS stack (large enough for any tree: < 22 levels x 512 pushes). */
template <bool BIG, int MODE>
__global__ __launch_bounds__(BIG ? 64 : 64 * GD_FW_WAVES) void k_walk_fast(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float4* __restrict__ wquad,
    const gd_tree_info* __restrict__ info, float3* __restrict__ acc, float* __restrict__ phi,
    float eps2, float G, float theta, int n, unsigned long long* __restrict__ stats,
    int* __restrict__ overflowList, int* __restrict__ overflowCount, int* __restrict__ bigStacks,
    const gd_bf16x4* __restrict__ mlpW, const float* __restrict__ mlpB, float minRatio) {
    constexpr bool QUAD = MODE != 0;
    constexpr bool GLOBAL_STACK = BIG;
    constexpr int W = BIG ? 1 : GD_FW_WAVES;
    constexpr int QW = QUAD ? W : 1;
    constexpr int QN = QUAD ? GD_FW_QLIST : 1;
    __shared__ int sStack[W][BIG ? GD_FW_BIG_STACK : GD_FW_STACK];
    __shared__ float4 sList[W][GD_FW_LIST];
    __shared__ float4 sQc[QW][QN], sQ0[QW][QN], sQ1[QW][QN];
    const unsigned lane = gd_lane();
    const int wave = GLOBAL_STACK ? 0 : (int)(threadIdx.x >> 6);
    long long g;
    int* stack;
    int cap;
    if (BIG) {
        g = overflowList[blockIdx.x];
        stack = sStack[0];
        cap = GD_FW_BIG_STACK;
        (void)bigStacks;
    } else {
        g = (long long)blockIdx.x * GD_FW_WAVES + wave;
        stack = sStack[wave];
        cap = GD_FW_STACK;
    }
    float4* list = sList[wave];
    float4* qc = sQc[QUAD ? wave : 0];
    float4* q0 = sQ0[QUAD ? wave : 0];
    float4* q1 = sQ1[QUAD ? wave : 0];
    if (g * 64 >= n) return;   /* wave-uniform */
    const long long i = g * 64 + lane;
    const bool valid = i < n;
    const float4 p = pos[valid ? i : g * 64];
    const float px = p.x, py = p.y, pz = p.z;

    /* group bounding box */
    float lx = px, ly = py, lz = pz, hx = px, hy = py, hz = pz;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lx = fminf(lx, __shfl_xor(lx, off)); hx = fmaxf(hx, __shfl_xor(hx, off));
        ly = fminf(ly, __shfl_xor(ly, off)); hy = fmaxf(hy, __shfl_xor(hy, off));
        lz = fminf(lz, __shfl_xor(lz, off)); hz = fmaxf(hz, __shfl_xor(hz, off));
    }
    const float ext = info->box.extent;
    const float th2 = theta * theta;
    const unsigned long long lt = gd_lanemask_lt();
    const float gcx = 0.5f * (lx + hx), gcy = 0.5f * (ly + hy), gcz = 0.5f * (lz + hz);

    gd_kahan3 tot = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float phs = 0.f, phc = 0.f;
    unsigned long long nInter = 0, nOpen = 0;
    int listCount = 0, qCount = 0;
    int top = 1;
    if (lane == 0) stack[0] = 0;
    bool overflowed = false;

    while (top > 0) {
        const int take = top < 64 ? top : 64;
        const int base = top - take;
        const bool have = (int)lane < take;
        int idx = 0;
        float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
        int2 lk = make_int2(0, 0);
        bool accept = false, open = false, leaf = false;
        float size = 0.f;
        if (have) {
            idx = stack[base + lane];
            c = wcom[idx];
            lk = wlink[idx];
            const int level = lk.y & 31;
            size = ldexpf(ext, -level);
            const float ddx = fmaxf(fmaxf(lx - c.x, c.x - hx), 0.f);
            const float ddy = fmaxf(fmaxf(ly - c.y, c.y - hy), 0.f);
            const float ddz = fmaxf(fmaxf(lz - c.z, c.z - hz), 0.f);
            const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            leaf = lk.x < 0;
            accept = leaf || (size * size < th2 * d2);
            open = !accept;
        }
        /* all of this wave's LDS reads of the popped entries are done before pushes */
        __builtin_amdgcn_wave_barrier();
        top = base;
        const bool toList = QUAD ? (accept && leaf) : accept;
        const unsigned long long accBal = __ballot(toList);
        const int nacc = __popcll(accBal);
        if (listCount + nacc > GD_FW_LIST) {
            float sx, sy, sz, sp;
            gd_eval_list(list, listCount, px, py, pz, eps2, sx, sy, sz, sp);
            gd_kahan_add(tot.x, tot.cx, sx);
            gd_kahan_add(tot.y, tot.cy, sy);
            gd_kahan_add(tot.z, tot.cz, sz);
            gd_kahan_add(phs, phc, sp);
            nInter += (unsigned long long)listCount;
            listCount = 0;
            __builtin_amdgcn_wave_barrier();
        }
        if (toList) list[listCount + __popcll(accBal & lt)] = c;
        listCount += nacc;
        if (QUAD) {
            const bool toQ = accept && !leaf;
            const unsigned long long qBal = __ballot(toQ);
            const int nq = __popcll(qBal);
            if (qCount + nq > GD_FW_QLIST) {
                float sx, sy, sz, sp;
                if (MODE == 2) {
                    float cx, cy, cz;
                    gd_eval_list(qc, qCount, px, py, pz, eps2, sx, sy, sz, sp);
                    gd_mlp_correct(qc, q0, q1, qCount, gcx, gcy, gcz, mlpW, mlpB, minRatio, cx,
                                   cy, cz);
                    sx += cx; sy += cy; sz += cz;
                } else {
                    gd_eval_quad(qc, q0, q1, qCount, px, py, pz, eps2, sx, sy, sz, sp);
                }
                gd_kahan_add(tot.x, tot.cx, sx);
                gd_kahan_add(tot.y, tot.cy, sy);
                gd_kahan_add(tot.z, tot.cz, sz);
                gd_kahan_add(phs, phc, sp);
                nInter += (unsigned long long)qCount;
                qCount = 0;
                __builtin_amdgcn_wave_barrier();
            }
            if (toQ) {
                const int at = qCount + __popcll(qBal & lt);
                qc[at] = c;
                q0[at] = wquad[2 * idx];
                float4 b1 = wquad[2 * idx + 1];
                if (MODE == 2) b1.w = size;
                q1[at] = b1;
            }
            qCount += nq;
        }
        /* push children of opened nodes */
        const int nk = open ? (lk.y >> 5) : 0;
        int incl = nk;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(incl, off);
            if ((int)lane >= off) incl += o;
        }
        const int total = __shfl(incl, 63);
        nOpen += (unsigned long long)__popcll(__ballot(open));
        if (top + total > cap) {
            overflowed = true;
            break;
        }
        if (open) {
            const int at = top + incl - nk;
            for (int k = 0; k < nk; ++k) stack[at + k] = lk.x + k;
        }
        top += total;
        __builtin_amdgcn_wave_barrier();
    }
    if (overflowed) {
        if (lane == 0) {
            if (GLOBAL_STACK) {
                atomicOr(&stats[2], 1ull);   /* reported as an error by the host */
            } else {
                const int slot = atomicAdd(overflowCount, 1);
                overflowList[slot] = (int)g;
            }
        }
        return;
    }
    if (listCount > 0) {
        float sx, sy, sz, sp;
        gd_eval_list(list, listCount, px, py, pz, eps2, sx, sy, sz, sp);
        gd_kahan_add(tot.x, tot.cx, sx);
        gd_kahan_add(tot.y, tot.cy, sy);
        gd_kahan_add(tot.z, tot.cz, sz);
        gd_kahan_add(phs, phc, sp);
        nInter += (unsigned long long)listCount;
    }
    if (QUAD && qCount > 0) {
        float sx, sy, sz, sp;
        if (MODE == 2) {
            float cx, cy, cz;
            gd_eval_list(qc, qCount, px, py, pz, eps2, sx, sy, sz, sp);
            gd_mlp_correct(qc, q0, q1, qCount, gcx, gcy, gcz, mlpW, mlpB, minRatio, cx, cy, cz);
            sx += cx; sy += cy; sz += cz;
        } else {
            gd_eval_quad(qc, q0, q1, qCount, px, py, pz, eps2, sx, sy, sz, sp);
        }
        gd_kahan_add(tot.x, tot.cx, sx);
        gd_kahan_add(tot.y, tot.cy, sy);
        gd_kahan_add(tot.z, tot.cz, sz);
        gd_kahan_add(phs, phc, sp);
        nInter += (unsigned long long)qCount;
    }
    if (valid) {
        acc[i] = make_float3(G * tot.x, G * tot.y, G * tot.z);
        if (phi) phi[i] = -G * phs;
    }
    if (stats && lane == 0) {
        const int nvalid = (int)((n - g * 64) < 64 ? (n - g * 64) : 64);
        atomicAdd(&stats[0], nInter * (unsigned long long)nvalid);
        atomicAdd(&stats[1], nOpen);
    }
}

/* ---------------------------------------------------------------------------
 * Production walk, split form: a traversal kernel writes each 64-body group's
 * interaction list to HBM, an evaluation kernel streams it.
 *
 * List layout: group g owns entries [g*cap, (g+1)*cap) of `lists`; accepted
 * leaves (and, in monopole mode, every accepted node) grow from the front,
 * accepted cells in quadrupole / learned mode grow from the back; counts[g] =
 * (front, back), or (-1, -1) when the group did not fit (its list or its LDS
 * stack) and is left to the fused kernel with a global-memory stack.
 * ------------------------------------------------------------------------- */
#ifndef GD_TR_STACK
#define GD_TR_STACK 1024          /* LDS stack entries per wavefront in the traversal */
#endif
#define GD_TR_WAVES 4
#define GD_EV_WAVES 4

template <bool SPLIT>
__global__ __launch_bounds__(64 * GD_TR_WAVES) void k_walk_traverse(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const gd_tree_info* __restrict__ info, float theta, int n,
    int cap, int* __restrict__ lists, int2* __restrict__ counts,
    int* __restrict__ overflowList, int* __restrict__ overflowCount,
    unsigned long long* __restrict__ stats) {
    __shared__ int sStack[GD_TR_WAVES][GD_TR_STACK];
    const unsigned lane = gd_lane();
    const int wave = (int)(threadIdx.x >> 6);
    const long long g = (long long)blockIdx.x * GD_TR_WAVES + wave;
    if (g * 64 >= n) return;   /* wave-uniform */
    int* stack = sStack[wave];
    int* list = lists + (size_t)g * cap;
    const long long i = g * 64 + lane;
    const float4 p = pos[i < n ? i : g * 64];
    float lx = p.x, ly = p.y, lz = p.z, hx = p.x, hy = p.y, hz = p.z;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lx = fminf(lx, __shfl_xor(lx, off)); hx = fmaxf(hx, __shfl_xor(hx, off));
        ly = fminf(ly, __shfl_xor(ly, off)); hy = fmaxf(hy, __shfl_xor(hy, off));
        lz = fminf(lz, __shfl_xor(lz, off)); hz = fmaxf(hz, __shfl_xor(hz, off));
    }
    const float ext = info->box.extent;
    const float th2 = theta * theta;
    const unsigned long long lt = gd_lanemask_lt();
    int nFront = 0, nBack = 0, top = 1;
    unsigned long long nOpen = 0;
    if (lane == 0) stack[0] = 0;
    bool overflowed = false;
    while (top > 0) {
        const int take = top < 64 ? top : 64;
        const int base = top - take;
        bool accept = false, open = false, leaf = false;
        int idx = 0, nk = 0, first = 0;
        if ((int)lane < take) {
            idx = stack[base + lane];
            const float4 c = wcom[idx];
            const int2 lk = wlink[idx];
            const float size = ldexpf(ext, -(lk.y & 31));
            const float ddx = fmaxf(fmaxf(lx - c.x, c.x - hx), 0.f);
            const float ddy = fmaxf(fmaxf(ly - c.y, c.y - hy), 0.f);
            const float ddz = fmaxf(fmaxf(lz - c.z, c.z - hz), 0.f);
            const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            leaf = lk.x < 0;
            accept = leaf || (size * size < th2 * d2);
            open = !accept;
            nk = open ? (lk.y >> 5) : 0;
            first = lk.x;
        }
        __builtin_amdgcn_wave_barrier();
        top = base;
        const bool front = SPLIT ? (accept && leaf) : accept;
        const bool back = SPLIT && accept && !leaf;
        const unsigned long long fb = __ballot(front);
        const unsigned long long bb = __ballot(back);
        const int nf = __popcll(fb), nb = __popcll(bb);
        /* children offsets: prefix sum of nk (0..8) from four bit-plane ballots */
        int excl = 0, total = 0;
#pragma unroll
        for (int bit = 0; bit < 4; ++bit) {
            const unsigned long long bal = __ballot((nk >> bit) & 1);
            excl += __popcll(bal & lt) << bit;
            total += __popcll(bal) << bit;
        }
        nOpen += (unsigned long long)__popcll(__ballot(open));
        if (nFront + nf + nBack + nb > cap || top + total > GD_TR_STACK) {
            overflowed = true;
            break;
        }
        if (front) list[nFront + __popcll(fb & lt)] = idx;
        if (back) list[cap - 1 - (nBack + __popcll(bb & lt))] = idx;
        nFront += nf;
        nBack += nb;
        if (open) {
            const int at = top + excl;
            for (int k = 0; k < nk; ++k) stack[at + k] = first + k;
        }
        top += total;
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) {
        if (overflowed) {
            counts[g] = make_int2(-1, -1);
            const int slot = atomicAdd(overflowCount, 1);
            overflowList[slot] = (int)g;
        } else {
            counts[g] = make_int2(nFront, nBack);
            if (stats) {
                const int nvalid = (int)((n - g * 64) < 64 ? (n - g * 64) : 64);
                atomicAdd(&stats[0], (unsigned long long)(nFront + nBack) * (unsigned long long)nvalid);
                atomicAdd(&stats[1], nOpen);
                atomicAdd(&stats[3], (unsigned long long)nBack * (unsigned long long)nvalid);
            }
        }
    }
}

/* 64 monopoles from LDS (zero-mass padding contributes nothing). */
template <bool PHI>
__device__ __forceinline__ void gd_eval64(const float4* __restrict__ e, float px, float py,
                                          float pz, float eps2, float& ax, float& ay, float& az,
                                          float& ph) {
#pragma unroll 8
    for (int k = 0; k < 64; ++k) {
        const float4 c = e[k];
        const float dx = c.x - px, dy = c.y - py, dz = c.z - pz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float mi = c.w * inv;
        const float f = mi * (inv * inv);
        ax = __builtin_fmaf(f, dx, ax);
        ay = __builtin_fmaf(f, dy, ay);
        az = __builtin_fmaf(f, dz, az);
        if (PHI) ph += mi;
    }
}

/* 64 quadrupole cells (com/mass in e, moments in q0/q1). */
template <bool PHI>
__device__ __forceinline__ void gd_eval64_quad(const float4* __restrict__ e,
                                               const float4* __restrict__ q0,
                                               const float4* __restrict__ q1, int cnt, float px,
                                               float py, float pz, float eps2, float& ax,
                                               float& ay, float& az, float& ph) {
    for (int k = 0; k < cnt; ++k) {
        const float4 c = e[k];
        const float4 a = q0[k];
        const float4 b = q1[k];
        const float dx = c.x - px, dy = c.y - py, dz = c.z - pz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv2 = inv * inv;
        const float mi = c.w * inv;
        const float inv5 = inv * inv2 * inv2;
        const float sdx = __builtin_fmaf(a.x, dx, __builtin_fmaf(a.w, dy, b.x * dz));
        const float sdy = __builtin_fmaf(a.w, dx, __builtin_fmaf(a.y, dy, b.y * dz));
        const float sdz = __builtin_fmaf(b.x, dx, __builtin_fmaf(b.y, dy, a.z * dz));
        const float q = __builtin_fmaf(dx, sdx, __builtin_fmaf(dy, sdy, dz * sdz));
        const float cd = __builtin_fmaf(mi, inv2, inv5 * __builtin_fmaf(7.5f * q, inv2, -1.5f * b.z));
        const float cs = -3.f * inv5;
        ax = __builtin_fmaf(cd, dx, __builtin_fmaf(cs, sdx, ax));
        ay = __builtin_fmaf(cd, dy, __builtin_fmaf(cs, sdy, ay));
        az = __builtin_fmaf(cd, dz, __builtin_fmaf(cs, sdz, az));
        if (PHI) ph += __builtin_fmaf(0.5f * inv5, __builtin_fmaf(3.f, q, -r2 * b.z), mi);
    }
}

/*
 * Evaluation: one wavefront per 64-body group streams its list in chunks of
 * 64 entries: each lane gathers one entry's record into LDS (double
 * buffered, the next chunk's gather is in flight while the current one is
 * evaluated), then all 64 bodies evaluate the chunk with broadcast LDS reads.
 * Each chunk's partial sum is added with Kahan compensation.  MODE 0: every
 * entry is a monopole; 1: back entries are quadrupole cells; 2: back entries
 * are quadrupole cells plus the network's residual correction (evaluated per
 * chunk at the group's centre on the matrix cores, scaled by each body's own
 * m s^2 / r^4).
 */
template <int MODE, bool PHI>
__global__ __launch_bounds__(64 * GD_EV_WAVES) void k_walk_eval(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float4* __restrict__ wquad,
    const gd_tree_info* __restrict__ info, const int* __restrict__ lists,
    const int2* __restrict__ counts, int cap, float3* __restrict__ acc,
    float* __restrict__ phi, float eps2, float G, int n, const gd_bf16x4* __restrict__ mlpW,
    const float* __restrict__ mlpB, float minRatio) {
    __shared__ float4 sE[GD_EV_WAVES][2][64];
    __shared__ float4 sQ0[MODE != 0 ? GD_EV_WAVES : 1][64];
    __shared__ float4 sQ1[MODE != 0 ? GD_EV_WAVES : 1][64];
    __shared__ float4 sCor[MODE == 2 ? GD_EV_WAVES : 1][64];
    const unsigned lane = gd_lane();
    const int wave = (int)(threadIdx.x >> 6);
    const long long g = (long long)blockIdx.x * GD_EV_WAVES + wave;
    if (g * 64 >= n) return;
    const int2 cnt = counts[g];
    if (cnt.x < 0) return;   /* walked by the fallback */
    const int* list = lists + (size_t)g * cap;
    const long long i = g * 64 + lane;
    const bool valid = i < n;
    const float4 p = pos[valid ? i : g * 64];
    const float px = p.x, py = p.y, pz = p.z;
    gd_kahan3 tot = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float phs = 0.f, phc = 0.f;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    /* padding of a partial chunk: massless and far away, contributes exactly 0 */
    const float4 pad = make_float4(1e18f, 1e18f, 1e18f, 0.f);

    /* front list: monopoles, double-buffered gathers */
    const int nf = cnt.x;
    float4 nxt = pad;
    if ((int)lane < nf) nxt = wcom[list[lane]];
    int buf = 0;
    for (int c0 = 0; c0 < nf; c0 += 64) {
        sE[wave][buf][lane] = nxt;
        __builtin_amdgcn_wave_barrier();
        nxt = pad;
        if (c0 + 64 + (int)lane < nf) nxt = wcom[list[c0 + 64 + lane]];
        float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
        gd_eval64<PHI>(sE[wave][buf], px, py, pz, eps2, ax, ay, az, ph);
        gd_kahan_add(tot.x, tot.cx, ax);
        gd_kahan_add(tot.y, tot.cy, ay);
        gd_kahan_add(tot.z, tot.cz, az);
        if (PHI) gd_kahan_add(phs, phc, ph);
        buf ^= 1;
    }
    /* back list: cells */
    if (MODE != 0) {
        const int nb = cnt.y;
        float gx = 0.f, gy = 0.f, gz = 0.f;
        if (MODE == 2) {
            float lx = px, ly = py, lz = pz, hx = px, hy = py, hz = pz;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                lx = fminf(lx, __shfl_xor(lx, off)); hx = fmaxf(hx, __shfl_xor(hx, off));
                ly = fminf(ly, __shfl_xor(ly, off)); hy = fmaxf(hy, __shfl_xor(hy, off));
                lz = fminf(lz, __shfl_xor(lz, off)); hz = fmaxf(hz, __shfl_xor(hz, off));
            }
            gx = 0.5f * (lx + hx); gy = 0.5f * (ly + hy); gz = 0.5f * (lz + hz);
        }
        const float ext = info->box.extent;
        for (int c0 = 0; c0 < nb; c0 += 64) {
            const int e = c0 + (int)lane;
            float4 c = zero4, q0 = zero4, q1 = zero4;
            int idx = -1;
            if (e < nb) {
                idx = list[cap - 1 - e];
                c = wcom[idx];
                q0 = wquad[2 * idx];
                q1 = wquad[2 * idx + 1];
            }
            const int m = nb - c0 < 64 ? nb - c0 : 64;
            float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
            if (MODE == 1) {
                sE[wave][0][lane] = c;
                sQ0[MODE != 0 ? wave : 0][lane] = q0;
                sQ1[MODE != 0 ? wave : 0][lane] = q1;
                __builtin_amdgcn_wave_barrier();
                gd_eval64_quad<PHI>(sE[wave][0], sQ0[MODE != 0 ? wave : 0],
                                    sQ1[MODE != 0 ? wave : 0], m, px, py, pz, eps2, ax, ay, az, ph);
            } else {
                /* network at the group centre, 16 cells per pass; lane l ends up
                   holding the correction (3 values) and m s^2 of cell l */
                float k = 0.f;
                float ox = 0.f, oy = 0.f, oz = 0.f;
                {
                    const float size = idx >= 0 ? ldexpf(ext, -(wlink[idx].y & 31)) : 1.f;
                    k = c.w * size * size;
                    const float dx = c.x - gx, dy = c.y - gy, dz = c.z - gz;
                    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                    const float rinv = idx >= 0 ? __builtin_amdgcn_rsqf(r2) : 0.f;
                    const float ux = dx * rinv, uy = dy * rinv, uz = dz * rinv;
                    const float ratio = size * rinv;
                    const float kinv = (idx >= 0 && k > 0.f) ? 1.f / k : 0.f;
                    const float su0 = __builtin_fmaf(q0.x, ux, __builtin_fmaf(q0.w, uy, q1.x * uz)) * kinv;
                    const float su1 = __builtin_fmaf(q0.w, ux, __builtin_fmaf(q0.y, uy, q1.y * uz)) * kinv;
                    const float su2 = __builtin_fmaf(q1.x, ux, __builtin_fmaf(q1.y, uy, q0.z * uz)) * kinv;
                    const float tr = q1.z * kinv;
                    const bool use = idx >= 0 && k > 0.f && ratio > minRatio;
                    /* pass t handles cells 16t..16t+15: features of cell (16t + l%16),
                       the feature quarter chosen by l/16, are fetched by lane shuffles */
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (16 * t >= m) break;
                        const int src = 16 * t + (int)(lane % 16);
                        const unsigned sel = lane / 16;
                        /* both feature quarters are fetched; the reader keeps its own */
                        const float u0 = __shfl(ux, src), s0 = __shfl(su0, src);
                        const float u1 = __shfl(uy, src), s1 = __shfl(su1, src);
                        const float u2 = __shfl(uz, src), s2 = __shfl(su2, src);
                        const float u3 = __shfl(ratio, src), s3 = __shfl(tr, src);
                        const float a0 = sel == 0 ? u0 : s0;
                        const float a1 = sel == 0 ? u1 : s1;
                        const float a2 = sel == 0 ? u2 : s2;
                        const float a3 = sel == 0 ? u3 : s3;
                        const gd_bf16x4 x = sel < 2 ? gd_pack4(a0, a1, a2, a3)
                                                    : gd_pack4(0.f, 0.f, 0.f, 0.f);
                        const gd_f32x4 o = gd_mlp_infer16(mlpW, mlpB, x);
                        /* outputs of cell 16t + j live in lane j (j < 16) */
                        const float r0 = __shfl(o[0], (int)(lane % 16));
                        const float r1 = __shfl(o[1], (int)(lane % 16));
                        const float r2o = __shfl(o[2], (int)(lane % 16));
                        if ((int)(lane / 16) == t) {
                            ox = r0; oy = r1; oz = r2o;
                        }
                    }
                    if (!use) k = 0.f;
                }
                sE[wave][0][lane] = c;
                sQ0[MODE != 0 ? wave : 0][lane] = q0;
                sQ1[MODE != 0 ? wave : 0][lane] = q1;
                sCor[MODE == 2 ? wave : 0][lane] = make_float4(ox, oy, oz, k);
                __builtin_amdgcn_wave_barrier();
                /* quadrupole cells plus the learned residual k / r^4 * o */
                gd_eval64_quad<PHI>(sE[wave][0], sQ0[MODE != 0 ? wave : 0],
                                    sQ1[MODE != 0 ? wave : 0], m, px, py, pz, eps2, ax, ay, az, ph);
                const float4* ce = sE[wave][0];
                const float4* co = sCor[MODE == 2 ? wave : 0];
                for (int kk = 0; kk < m; ++kk) {
                    const float4 cc = ce[kk];
                    const float4 oo = co[kk];
                    const float dx = cc.x - px, dy = cc.y - py, dz = cc.z - pz;
                    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                    const float i2 = __builtin_amdgcn_rcpf(r2);
                    const float s4 = oo.w * (i2 * i2);
                    ax = __builtin_fmaf(s4, oo.x, ax);
                    ay = __builtin_fmaf(s4, oo.y, ay);
                    az = __builtin_fmaf(s4, oo.z, az);
                }
            }
            gd_kahan_add(tot.x, tot.cx, ax);
            gd_kahan_add(tot.y, tot.cy, ay);
            gd_kahan_add(tot.z, tot.cz, az);
            if (PHI) gd_kahan_add(phs, phc, ph);
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (valid) {
        acc[i] = make_float3(G * tot.x, G * tot.y, G * tot.z);
        if (PHI) phi[i] = -G * phs;
    }
}

/* Walk scratch: stats[4], overflow count, overflow list, big stacks. */
struct gd_walk_scratch {
    unsigned long long* stats;
    int* overflowCount;
    int* overflowList;
    int* bigStacks;
    int listCap;
    int bigSlots;
    int* lists;            /* groups x entriesPerGroup interaction lists */
    int2* counts;          /* per group (front, back) */
    long long listBytes;
    int entriesPerGroup;
};

static gd_walk_scratch g_walk = {nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, 0, 16384};

/* optional HIP-event timing of the walk kernel (null stream) */
static int g_walkTiming = 0;
static hipEvent_t g_walkEv[4] = {nullptr, nullptr, nullptr, nullptr};
static int g_walkTimed = 0;

static void gd_walk_ensure(int groups) {
    if (!g_walk.stats) {
        GD_HIP(hipMalloc(&g_walk.stats, 4 * sizeof(unsigned long long)));
        GD_HIP(hipMalloc(&g_walk.overflowCount, sizeof(int)));
        GD_HIP(hipMemset(g_walk.stats, 0, 4 * sizeof(unsigned long long)));
        GD_HIP(hipMemset(g_walk.overflowCount, 0, sizeof(int)));
    }
    if (groups > g_walk.listCap) {
        if (g_walk.overflowList) GD_HIP(hipFree(g_walk.overflowList));
        if (g_walk.counts) GD_HIP(hipFree(g_walk.counts));
        GD_HIP(hipMalloc(&g_walk.overflowList, sizeof(int) * (size_t)groups));
        GD_HIP(hipMalloc(&g_walk.counts, sizeof(int2) * (size_t)groups));
        g_walk.listCap = groups;
    }
    const long long need = (long long)groups * g_walk.entriesPerGroup * (long long)sizeof(int);
    if (need > g_walk.listBytes) {
        if (g_walk.lists) GD_HIP(hipFree(g_walk.lists));
        GD_HIP(hipMalloc(&g_walk.lists, (size_t)need));
        g_walk.listBytes = need;
    }
}

__global__ void k_walk_reset(unsigned long long* stats, int* overflowCount) {
    if (threadIdx.x < 4) stats[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) *overflowCount = 0;
}

static void gd_walk_impl(const treeState* tree, const float4* pos, float3* acc, float* phi,
                         float eps2, float G, float theta, int starCount, gd_walk_stats* stats,
                         int mode, const gd_mlp* net, float minRatio) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_walk: tree not initialised");
        return;
    }
    if (starCount <= 0) return;
    gd_builder* b = gd_builder_of(tree->builder);
    if (mode != 0 && !(b->quad && b->wquad)) {
        GD_FAIL("gd_walk: this walk needs the second moments (gd_set_quadrupole before the build)");
        return;
    }
    const int groups = gd_div_up(starCount, 64);
    gd_walk_ensure(groups);
    if (gd_has_error()) return;
    const gd_bf16x4* W = net ? gd_mlp_fwd_tiles(net) : nullptr;
    const float* Bv = net ? gd_mlp_bias(net) : nullptr;
    k_walk_reset<<<1, 64>>>(g_walk.stats, g_walk.overflowCount);
    GD_LAUNCHED();
    if (g_walkTiming) {
        if (!g_walkEv[0])
            for (int e = 0; e < 4; ++e) GD_HIP(hipEventCreate(&g_walkEv[e]));
        GD_HIP(hipEventRecord(g_walkEv[0], nullptr));
    }
#define GD_WALK_ARGS(LIST, STACKS)                                                             \
    pos, b->wcom, b->wlink, b->wquad, b->dinfo, acc, phi, eps2, G, theta, starCount,           \
        g_walk.stats, LIST, g_walk.overflowCount, STACKS, W, Bv, minRatio
    const int cap = g_walk.entriesPerGroup;
    const int tblocks = gd_div_up(groups, GD_TR_WAVES);
    if (mode == 0)
        k_walk_traverse<false><<<tblocks, 64 * GD_TR_WAVES>>>(
            pos, b->wcom, b->wlink, b->dinfo, theta, starCount, cap, g_walk.lists, g_walk.counts,
            g_walk.overflowList, g_walk.overflowCount, g_walk.stats);
    else
        k_walk_traverse<true><<<tblocks, 64 * GD_TR_WAVES>>>(
            pos, b->wcom, b->wlink, b->dinfo, theta, starCount, cap, g_walk.lists, g_walk.counts,
            g_walk.overflowList, g_walk.overflowCount, g_walk.stats);
    GD_LAUNCHED();
    if (g_walkTiming) GD_HIP(hipEventRecord(g_walkEv[1], nullptr));
    const int eblocks = gd_div_up(groups, GD_EV_WAVES);
#define GD_EVAL_ARGS                                                                           \
    pos, b->wcom, b->wlink, b->wquad, b->dinfo, g_walk.lists, g_walk.counts, cap, acc, phi,    \
        eps2, G, starCount, W, Bv, minRatio
    if (mode == 0) {
        if (phi) k_walk_eval<0, true><<<eblocks, 64 * GD_EV_WAVES>>>(GD_EVAL_ARGS);
        else k_walk_eval<0, false><<<eblocks, 64 * GD_EV_WAVES>>>(GD_EVAL_ARGS);
    } else if (mode == 1) {
        if (phi) k_walk_eval<1, true><<<eblocks, 64 * GD_EV_WAVES>>>(GD_EVAL_ARGS);
        else k_walk_eval<1, false><<<eblocks, 64 * GD_EV_WAVES>>>(GD_EVAL_ARGS);
    } else {
        if (phi) k_walk_eval<2, true><<<eblocks, 64 * GD_EV_WAVES>>>(GD_EVAL_ARGS);
        else k_walk_eval<2, false><<<eblocks, 64 * GD_EV_WAVES>>>(GD_EVAL_ARGS);
    }
    GD_LAUNCHED();
    if (g_walkTiming) GD_HIP(hipEventRecord(g_walkEv[2], nullptr));
#undef GD_EVAL_ARGS
    /* groups whose list or stack did not fit: the fused walk with a global stack */
    int over = 0;
    GD_HIP(hipMemcpy(&over, g_walk.overflowCount, sizeof(int), hipMemcpyDeviceToHost));
    if (over > 0) {
        if (mode == 0)
            k_walk_fast<true, 0><<<over, 64>>>(GD_WALK_ARGS(g_walk.overflowList, nullptr));
        else if (mode == 1)
            k_walk_fast<true, 1><<<over, 64>>>(GD_WALK_ARGS(g_walk.overflowList, nullptr));
        else
            k_walk_fast<true, 2><<<over, 64>>>(GD_WALK_ARGS(g_walk.overflowList, nullptr));
        GD_LAUNCHED();
    }
#undef GD_WALK_ARGS
    if (g_walkTiming) {
        GD_HIP(hipEventRecord(g_walkEv[3], nullptr));
        g_walkTimed = 1;
    }
    /* more than 1 group in 256 overflowed its list: give later walks room */
    if (over > groups / 256 && g_walk.entriesPerGroup < 65536) g_walk.entriesPerGroup *= 2;
    if (over > 0 || stats) {
        unsigned long long h[4];
        GD_HIP(hipMemcpy(h, g_walk.stats, sizeof(h), hipMemcpyDeviceToHost));
        if (h[2]) GD_FAIL("gd_walk: a group's traversal stack exceeded %d entries", GD_FW_BIG_STACK);
        if (!stats) return;
        stats->interactions = h[0];
        stats->opened = h[1];
        stats->overflowGroups = (unsigned long long)over;
        stats->cellInteractions = h[3];
    }
}

extern "C" void gd_walk(const treeState* tree, const float4* pos, float3* acc, float* phi,
                        float eps2, float G, float theta, int starCount, gd_walk_stats* stats) {
    const gd_builder* b = tree && tree->builder ? gd_builder_of(tree->builder) : nullptr;
    const int mode = (b && b->quad && b->wquad) ? 1 : 0;
    gd_walk_impl(tree, pos, acc, phi, eps2, G, theta, starCount, stats, mode, nullptr, 0.f);
}

extern "C" void gd_walk_learned(const treeState* tree, const float4* pos, float3* acc,
                                float* phi, float eps2, float G, float theta, int starCount,
                                const gd_mlp* net, float minRatio, gd_walk_stats* stats) {
    if (!net) {
        GD_FAIL("gd_walk_learned: no network");
        return;
    }
    gd_walk_impl(tree, pos, acc, phi, eps2, G, theta, starCount, stats, 2, net, minRatio);
}

extern "C" void gd_walk_timing(int enable) { g_walkTiming = enable ? 1 : 0; }

extern "C" float gd_walk_kernel_ms(void) {
    float ms[3];
    if (gd_walk_phase_ms(ms) != 0) return -1.0f;
    return ms[1];
}

extern "C" int gd_walk_phase_ms(float* ms) {
    if (!g_walkTimed) return -1;
    GD_HIP(hipEventSynchronize(g_walkEv[3]));
    for (int e = 0; e < 3; ++e) GD_HIP(hipEventElapsedTime(&ms[e], g_walkEv[e], g_walkEv[e + 1]));
    return 0;
}
