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
 * 2. Production walk (gd_walk, gd_walk_learned).  Bodies are taken in
 *    64-body groups (consecutive in Morton order), one group per wavefront.
 *    A traversal kernel tests the group's bounding box against 64 nodes at a
 *    time, one node per lane: a node is accepted when it is a leaf or
 *    size^2 < theta^2 d^2, d being the distance from the group box to the
 *    node's centre of mass and size = bbox cube extent / 2^level; opened
 *    nodes push their (contiguous) children on an LDS stack.  Accepted node
 *    indices are written to the group's interaction list in HBM.  An
 *    evaluation kernel then streams each list in 64-entry chunks through LDS
 *    and every body of the group evaluates every entry (compute bound: ~20
 *    FP32 FLOP per monopole, ~54 per quadrupole cell).  Groups whose list or
 *    stack does not fit are walked body by body (k_walk_body).
 */
#include "gd_internal.h"
#include "gd_far.h"
#include "gd_mlp.h"
#include "gd_traverse.h"

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
    k_walk_parity<<<gd_div_up(groups, GD_PAR_SLOTS), 256, 0, gd_stream()>>>(
        positionMassArray, tree.nodes, tree.boundingData[2], tree.boundingData[3], accelVals,
        antiSingularitySquared, GVal, theta, starCount);
    GD_LAUNCHED();
}

#define GD_BODY_STACK 12288       /* per-body walk LDS stack (48 KB): > 22 levels x 512 pushes */
#define GD_BODY_WAVES 4           /* per-body walk: bodies (wavefronts) per block, first try */

struct gd_kahan3 {
    float x, y, z, cx, cy, cz;
};

__device__ __forceinline__ void gd_kahan_add(float& s, float& c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

/* ---------------------------------------------------------------------------
 * Production walk: a traversal kernel writes each 64-body group's interaction
 * list to HBM, an evaluation kernel streams it.
 *
 * List layout: group g owns entries [g*cap, (g+1)*cap) of `lists`; accepted
 * leaves (and, in monopole mode, every accepted node) grow from the front,
 * accepted cells in quadrupole / learned mode grow from the back; counts[g] =
 * (front, back), or (-1, -1) when the group did not fit (its list or its LDS
 * stack) and is left to the per-body walk.
 * ------------------------------------------------------------------------- */
#ifndef GD_TR_STACK
#define GD_TR_STACK 1024          /* LDS stack ranges per wavefront, first pass: with the
                                     pop buffer, 4640 bytes, eight wavefronts per SIMD */
#endif
#define GD_TR_WAVES 4
#define GD_EV_WAVES 4
#define GD_TR2_STACK 4096         /* second pass, one wavefront per block (16 KB of LDS) */
#define GD_POP (64 * GD_TR_PER + 8)   /* LDS pop buffer per traversing wavefront: the
                                        step's nodes and the 8-int slack of the range
                                        expansion */
#define GD_BODY_STACK1 2048       /* per-body walk, first try (nodes) */

/* Walk statistics: 64 copies of the counters, 128 bytes apart, picked by
 * block index, so that a launch's per-group atomics spread over 64 cache
 * lines instead of queueing on one; the host sums them.  [0] interactions,
 * [1] opened nodes, [2] per-body stack overflow flag, [3] cell interactions,
 * [4] network evaluations. */
#define GD_STAT_SLOTS 64
#define GD_STAT_STRIDE 16
#define GD_STAT_WORDS (GD_STAT_SLOTS * GD_STAT_STRIDE)
__device__ __forceinline__ unsigned long long* gd_stat_slot(unsigned long long* stats) {
    return stats + (blockIdx.x & (GD_STAT_SLOTS - 1)) * GD_STAT_STRIDE;
}
/* A launch's slot count: `slots`, or at most *slotsDev when the count is
 * held on the device (the walk's asynchronous mode sizes grids for a fixed
 * capacity and reads how many slots are in use there). */
__device__ __forceinline__ long long gd_slots(int slots, const int* slotsDev) {
    if (!slotsDev) return slots;
    const int d = *slotsDev;
    return d < slots ? d : slots;
}
/* Segments per second-pass list when `groups` lists are evaluated at once:
 * enough wavefronts to fill GD_SEG_WAVES, at most 64 each. */
__host__ __device__ __forceinline__ int gd_seg_count(int groups) {
    if (groups <= 0) return 1;
    if (groups > 2048 / 2) return 1;
    const int s = 2048 / groups;
    return s < 64 ? s : 64;
}
#define GD_LIST_MIN 4096          /* first-pass room of any group (entries) */
#define GD_PASS2_FLOOR 16384      /* the second pass never gives up on a shorter list */
#define GD_SPREAD (-2)            /* sizes[] mark: the group is walked body by body */

/*
 * Traversal kernel, one wavefront per group.  Slot s walks group
 * groupList[s] (or group s when groupList is NULL).  Its list is
 * lists[offsets[g] .. + caps[g]) for group g when offsets is given (the
 * first pass's packed layout), else lists[s*cap .. + cap); it gives up once the list
 * passes slotCap[s] (when given) or its capacity, or the LDS stack of STACK
 * entries overflows: counts = (-1, -1) and the group goes to
 * overflowList.  Otherwise counts = (front, back) -- counts[g] for group g
 * in the first pass (offsets given), counts[s] otherwise -- and, when sizes is
 * given, sizes[g] records the list length for the next walk's layout.  In
 * the second pass (offsets NULL) a group given up on is marked spread
 * (sizes -2); the first pass, when spreadList is given, sends a marked group
 * straight to the per-body walk (spreadList) without traversing it.
 */
template <bool SPLIT, int WAVES, int STACK>
__global__ __launch_bounds__(64 * WAVES) void k_walk_traverse(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const gd_tree_info* __restrict__ info, float theta, int n,
    const int2* __restrict__ slotRange, const int* __restrict__ groupList, int slots,
    const long long* __restrict__ offsets,
    const int* __restrict__ caps, int cap, const int* __restrict__ slotCap,
    int* __restrict__ lists, int2* __restrict__ counts, int* __restrict__ sizes,
    int* __restrict__ overflowList, int* __restrict__ overflowCount,
    int* __restrict__ spreadList, int* __restrict__ spreadCount, int* __restrict__ deepCount,
    unsigned long long* __restrict__ stats, const int* __restrict__ slotsDev = nullptr) {
    __shared__ int sStack[WAVES][STACK + GD_POP];
    if (info->overflow) return;   /* the build ran out of nodes: no tree to walk (the host
                                     reports it from the walk's status) */
    const int wave = WAVES > 1 ? (int)(threadIdx.x >> 6) : 0;
    const long long s = (long long)gd_xcd_block() * WAVES + wave;
    if (s >= gd_slots(slots, slotsDev)) return;   /* wave-uniform */
    const long long g = groupList ? groupList[s] : s;
    const int2 rg = slotRange[g];   /* (first body, bodies) */
    const long long cs = offsets ? g : s;   /* counts by slot id in the first pass */
    if (rg.y == 0) {                /* empty second slot of a whole run */
        if (gd_lane() == 0) counts[cs] = make_int2(0, 0);
        return;
    }
    if (spreadList && sizes[g] == GD_SPREAD) {
        if (gd_lane() == 0) {
            counts[cs] = make_int2(-1, -1);
            spreadList[atomicAdd(spreadCount, 1)] = (int)g;
        }
        return;
    }
    float lx, ly, lz, hx, hy, hz;
    gd_range_box(pos, rg.x, rg.y, lx, ly, lz, hx, hy, hz);
    int nf = 0, nb = 0;
    unsigned long long nOpen = 0;
    int* list;
    int size;
    if (offsets) {
        list = lists + offsets[g];
        size = caps[g];
    } else {
        list = lists + (size_t)s * cap;
        size = cap;
    }
    const int limit = slotCap ? min(slotCap[s], size) : size;
    int depth = 0;
    const bool ok = gd_traverse_box<SPLIT>(wcom, wlink, wrad, info, theta, lx, ly, lz, hx, hy, hz,
                                           sStack[wave], STACK, sStack[wave] + STACK, list, size,
                                           limit, &nf, &nb,
                                           &nOpen, &depth);
    if (gd_lane() == 0) {
        /* groups deeper than 7/8 of GD_TR_STACK (or overflowing it) */
        if (deepCount && depth > GD_TR_STACK - GD_TR_STACK / 8) atomicAdd(deepCount, 1);
        if (ok) {
            counts[cs] = make_int2(nf, nb);
            if (sizes) sizes[g] = nf + nb;
            if (stats) {
                const int nvalid = rg.y;
                unsigned long long* sl = gd_stat_slot(stats);
                atomicAdd(&sl[0], (unsigned long long)(nf + nb) * (unsigned long long)nvalid);
                atomicAdd(&sl[1], nOpen);
                atomicAdd(&sl[3], (unsigned long long)nb * (unsigned long long)nvalid);
            }
        } else {
            counts[cs] = make_int2(-1, -1);
            if (sizes) sizes[g] = offsets ? -1 : GD_SPREAD;
            const int at = atomicAdd(overflowCount, 1);
            overflowList[at] = (int)g;
        }
    }
}

/*
 * First-pass list layout: group g gets room for one and a half times its
 * last list length, at least GD_LIST_MIN entries, rounded up to 64 entries;
 * a group marked spread (sizes -2, see k_walk_traverse) gets none -- or
 * every group gets exactly `fixed` entries when fixed > 0.  The rooms are
 * packed back to back by an exclusive scan (one 1024-thread block).  The
 * margin keeps groups whose list grows between steps (a Morton run that
 * shifts onto a neighbouring cluster) in the first pass, where their work
 * hides among the others'; a long list walked alone in the second pass would
 * hold up the end of the walk.  With `reset`, spread marks are cleared
 * (the group is tried again as a list).  Groups that would end
 * beyond `capacity` entries get no room (they go to the second pass);
 * *total receives the unclamped sum, for growing the pool.
 *
 * The same scan numbers the non-empty slots: order[] lists them in slot
 * (Morton) order, and its tail -- as many entries as there are empty slots
 * -- holds `groups`, the index of an always-empty sentinel slot.  The first
 * pass walks slot order[s] with wavefront s, so no wavefront of a busy block
 * idles on the empty second slot of a run that was not split.
 */
__device__ __forceinline__ long long gd_list_room(int sz, int fixed) {
    if (fixed > 0) return fixed;
    if (sz == GD_SPREAD) return 0;
    long long want = sz >= 0 ? sz + sz / 2 : GD_LIST_MIN;
    want = want > GD_LIST_MIN ? want : GD_LIST_MIN;
    return (want + 63) & ~63ll;
}

/* Block-wide exclusive scan of one 64-bit value per thread (1024 threads);
 * returns the thread's exclusive prefix, *blockTotal the block's sum. */
__device__ __forceinline__ long long gd_block_scan1024(long long v, long long* sWave,
                                                      long long* blockTotal) {
    long long inc = v;
    const unsigned lane = gd_lane();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long o = __shfl_up(inc, off);
        if ((int)lane >= off) inc += o;
    }
    if (lane == 63) sWave[threadIdx.x / 64] = inc;
    __syncthreads();
    long long base = 0, all = 0;
    for (int w = 0; w < 16; ++w) {
        const long long t = sWave[w];
        if (w < (int)(threadIdx.x / 64)) base += t;
        all += t;
    }
    *blockTotal = all;
    return base + inc - v;
}

/* Layout, part 1: each 1024-thread block scans the rooms of its 1024 groups
 * (one per thread, coalesced) into block-local offsets and its total. */
__global__ __launch_bounds__(1024) void k_walk_layout(const int2* __restrict__ slotRange,
                                                      int* __restrict__ sizes, int groups,
                                                      int fixed, int reset,
                                                      long long* __restrict__ offsets,
                                                      long long* __restrict__ blockSums) {
    __shared__ long long sWave[16];
    __shared__ long long sFlag[16];
    const int g = blockIdx.x * 1024 + threadIdx.x;
    long long room = 0, busy = 0;
    if (g < groups && slotRange[g].y > 0) {
        if (reset && sizes[g] == GD_SPREAD) sizes[g] = -1;
        room = gd_list_room(sizes[g], fixed);
        busy = 1;
    }
    long long all, slots;
    const long long at = gd_block_scan1024(room, sWave, &all);
    gd_block_scan1024(busy, sFlag, &slots);
    if (g < groups) offsets[g] = at;
    if (threadIdx.x == 0) {
        blockSums[blockIdx.x] = all;
        blockSums[gridDim.x + blockIdx.x] = slots;   /* non-empty slots */
    }
}

/* Layout, part 2: adds the preceding blocks' totals, gives each group its
 * room when it ends within `capacity` (else none), and the last block
 * writes the requested total; writes order[]. */
__global__ __launch_bounds__(1024) void k_walk_place(const int2* __restrict__ slotRange,
                                                     const int* __restrict__ sizes, int groups,
                                                     int fixed, long long capacity,
                                                     const long long* __restrict__ blockSums,
                                                     long long* __restrict__ offsets,
                                                     int* __restrict__ caps,
                                                     int* __restrict__ order,
                                                     unsigned long long* __restrict__ total) {
    __shared__ long long sBase, sSlotBase, sSlots;
    __shared__ long long sFlag[16];
    {
        /* the preceding blocks' rooms and non-empty slots, and all blocks'
           slots: block-wide sums (one term per thread) */
        long long base = 0, slotBase = 0, slots = 0;
        for (unsigned k = threadIdx.x; k < gridDim.x; k += 1024) {
            const long long c = blockSums[gridDim.x + k];
            if (k < blockIdx.x) {
                base += blockSums[k];
                slotBase += c;
            }
            slots += c;
        }
        long long t0, t1, t2;
        /* one shared buffer: every scan's reads end before the next one writes */
        gd_block_scan1024(base, sFlag, &t0);
        __syncthreads();
        gd_block_scan1024(slotBase, sFlag, &t1);
        __syncthreads();
        gd_block_scan1024(slots, sFlag, &t2);
        __syncthreads();
        if (threadIdx.x == 0) {
            sBase = t0;
            sSlotBase = t1;
            sSlots = t2;
            if (blockIdx.x == gridDim.x - 1)
                *total = (unsigned long long)(t0 + blockSums[blockIdx.x]);
        }
    }
    __syncthreads();
    const int g = blockIdx.x * 1024 + threadIdx.x;
    const bool busy = g < groups && slotRange[g].y > 0;
    long long blockSlots;
    const long long k = sSlotBase + gd_block_scan1024(busy ? 1 : 0, sFlag, &blockSlots);
    if (g >= groups) return;
    if (busy) order[k] = g;
    if (g >= sSlots) order[g] = groups;   /* the tail: the sentinel slot */
    const long long at = sBase + offsets[g];
    const long long room = slotRange[g].y > 0 ? gd_list_room(sizes[g], fixed) : 0;
    offsets[g] = at;
    caps[g] = at + room <= capacity ? (int)room : 0;
}

/*
 * Estimate for the groups the first pass could not list: the list length of
 * one of the group's bodies (the middle one in Morton order), counted by a
 * point traversal.  The second pass gives up on the group -- leaving it to
 * the per-body walk -- once its list passes GD_BODY_RATIO times that, and
 * at least `floor` entries: a compact group's list is 1.5-3x its bodies'
 * own, one that much longer is spread out (its box spans distant clusters).
 * Such a group is cheaper body by body -- the per-body walk does ~25x less
 * work per second than the list evaluation, so short lists stay lists -- and
 * a long list walked by a single wavefront would also hold up the end of the
 * walk.
 */
#define GD_BODY_RATIO 8
template <bool SPLIT>
__global__ __launch_bounds__(64 * GD_TR_WAVES) void k_walk_estimate(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const gd_tree_info* __restrict__ info, float theta, int n,
    const int2* __restrict__ slotRange, const int* __restrict__ groupList, int slots, int cap2,
    int floor, int* __restrict__ slotCap, const int* __restrict__ slotsDev = nullptr) {
    __shared__ int sStack[GD_TR_WAVES][GD_TR_STACK + GD_POP];
    const int wave = (int)(threadIdx.x >> 6);
    const long long s = (long long)blockIdx.x * GD_TR_WAVES + wave;
    if (s >= gd_slots(slots, slotsDev)) return;   /* wave-uniform */
    const int2 rg = slotRange[groupList[s]];
    const float4 p = pos[rg.x + (rg.y - 1) / 2];
    int nf = 0, nb = 0, depth = 0;
    unsigned long long nOpen = 0;
    const int bound = cap2 / GD_BODY_RATIO;
    const bool ok = gd_traverse_box<SPLIT>(wcom, wlink, wrad, info, theta, p.x, p.y, p.z, p.x, p.y,
                                           p.z, sStack[wave], GD_TR_STACK,
                                           sStack[wave] + GD_TR_STACK, nullptr, bound,
                                           bound, &nf, &nb, &nOpen, &depth);
    if (gd_lane() == 0) {
        long long lim = ok ? (long long)GD_BODY_RATIO * (nf + nb) : cap2;
        lim = lim > floor ? lim : floor;
        slotCap[s] = (int)(lim < cap2 ? lim : cap2);
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

/* 64 quadrupole cells: com/mass in e, second moments in q0/q1 prescaled for
 * this loop, 7.5 (Sxx, Syy, Szz, Sxy) and (7.5 Sxz, 7.5 Syz, -1.5 tr S, tr S)
 * (two multiplies per cell instead of per body and cell).  Padding entries
 * (massless, far away, zero moments) contribute exactly 0, so the loop always
 * runs 64 entries and needs no per-lane trip count. */
template <bool PHI>
__device__ __forceinline__ void gd_eval64_quad(const float4* __restrict__ e,
                                               const float4* __restrict__ q0,
                                               const float4* __restrict__ q1, float px,
                                               float py, float pz, float eps2, float& ax,
                                               float& ay, float& az, float& ph) {
#pragma unroll 2
    for (int k = 0; k < 64; ++k) {
        const float4 c = e[k];
        const float4 a = q0[k];
        const float4 b = q1[k];
        const float dx = c.x - px, dy = c.y - py, dz = c.z - pz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv2 = inv * inv;
        const float inv3 = inv * inv2;
        const float inv5 = inv3 * inv2;
        /* with S' = 7.5 S: sd = S' d, q = d.S' d */
        const float sdx = __builtin_fmaf(a.x, dx, __builtin_fmaf(a.w, dy, b.x * dz));
        const float sdy = __builtin_fmaf(a.w, dx, __builtin_fmaf(a.y, dy, b.y * dz));
        const float sdz = __builtin_fmaf(b.x, dx, __builtin_fmaf(b.y, dy, a.z * dz));
        const float q = __builtin_fmaf(dx, sdx, __builtin_fmaf(dy, sdy, dz * sdz));
        const float cd = __builtin_fmaf(c.w, inv3, inv5 * __builtin_fmaf(q, inv2, b.z));
        const float cs = -0.4f * inv5;
        ax = __builtin_fmaf(cd, dx, __builtin_fmaf(cs, sdx, ax));
        ay = __builtin_fmaf(cd, dy, __builtin_fmaf(cs, sdy, ay));
        az = __builtin_fmaf(cd, dz, __builtin_fmaf(cs, sdz, az));
        if (PHI) ph += __builtin_fmaf(0.5f * inv5, __builtin_fmaf(0.4f, q, -r2 * b.w), c.w * inv);
    }
}

/* The same for 64 cells stored as one record each, e3[3k .. 3k+2] = (com,
 * 7.5 S first four, 7.5 Sxz 7.5 Syz -1.5 trS trS): one LDS base address
 * serves all three reads of an entry. */
template <bool PHI>
__device__ __forceinline__ void gd_eval64_quad3(const float4* __restrict__ e3, float px, float py,
                                                float pz, float eps2, float& ax, float& ay,
                                                float& az, float& ph) {
#pragma unroll 2
    for (int k = 0; k < 64; ++k) {
        const float4 c = e3[3 * k];
        const float4 a = e3[3 * k + 1];
        const float4 b = e3[3 * k + 2];
        const float dx = c.x - px, dy = c.y - py, dz = c.z - pz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv2 = inv * inv;
        const float inv3 = inv * inv2;
        const float inv5 = inv3 * inv2;
        const float sdx = __builtin_fmaf(a.x, dx, __builtin_fmaf(a.w, dy, b.x * dz));
        const float sdy = __builtin_fmaf(a.w, dx, __builtin_fmaf(a.y, dy, b.y * dz));
        const float sdz = __builtin_fmaf(b.x, dx, __builtin_fmaf(b.y, dy, a.z * dz));
        const float q = __builtin_fmaf(dx, sdx, __builtin_fmaf(dy, sdy, dz * sdz));
        const float cd = __builtin_fmaf(c.w, inv3, inv5 * __builtin_fmaf(q, inv2, b.z));
        const float cs = -0.4f * inv5;
        ax = __builtin_fmaf(cd, dx, __builtin_fmaf(cs, sdx, ax));
        ay = __builtin_fmaf(cd, dy, __builtin_fmaf(cs, sdy, ay));
        az = __builtin_fmaf(cd, dz, __builtin_fmaf(cs, sdz, az));
        if (PHI) ph += __builtin_fmaf(0.5f * inv5, __builtin_fmaf(0.4f, q, -r2 * b.w), c.w * inv);
    }
}

/*
 * The network on the m cells of a compacted list: cell c (c < m) is held by
 * lane map[c] (map in LDS), whose 8 features f[0..7] are fetched by
 * shuffles.  Pass t evaluates cells 16t..16t+15 on the matrix cores (column j
 * of the 16x16 input tile is cell 16t + j; the lanes of rows 0-3 / 4-7 fetch
 * features 0-3 / 4-7).  On return lane c holds cell c's scaled residual
 * (ox, oy, oz).
 */
__device__ __forceinline__ void gd_mlp_cells(const gd_bf16x4* __restrict__ W,
                                             const float* __restrict__ Bv, int m,
                                             const int* map, const float* f, float& ox,
                                             float& oy, float& oz) {
    const unsigned lane = gd_lane();
    ox = oy = oz = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (16 * t >= m) break;   /* wave-uniform */
        const int c = 16 * t + (int)(lane % 16);
        const int src = c < m ? map[c] : 0;
        const unsigned sel = lane / 16;
        float a[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            /* both quarters are fetched; the reader keeps its own */
            const float lo = __shfl(f[q], src), hi = __shfl(f[4 + q], src);
            a[q] = sel == 0 ? lo : hi;
        }
        const gd_bf16x4 x = sel < 2 ? gd_pack4(a[0], a[1], a[2], a[3])
                                    : gd_pack4(0.f, 0.f, 0.f, 0.f);
        /* an offset the compiler cannot see through: the weight and bias reads
           stay in the pass instead of being hoisted into ~130 registers */
        int off = 0;
        asm volatile("" : "+v"(off));
        const gd_f32x4 o = gd_mlp_infer16(W + off, Bv + off, x);
        /* outputs of cell 16t + j live in lane j (j < 16) */
        const float r0 = __shfl(o[0], (int)(lane % 16));
        const float r1 = __shfl(o[1], (int)(lane % 16));
        const float r2 = __shfl(o[2], (int)(lane % 16));
        if ((int)(lane / 16) == t) {
            ox = r0; oy = r1; oz = r2;
        }
    }
}

/* The network for the lanes with need set, results compacted: returns their
 * number n (wave-uniform); for i < n, corr[i] = (k * (ox, oy, oz), the
 * lane's index as int bits) of the i-th such lane in lane order, with k that
 * lane's scale.  map: 64 ints of LDS; corr: 64 float4 of LDS. */
__device__ __forceinline__ int gd_mlp_compact(const gd_bf16x4* __restrict__ W,
                                             const float* __restrict__ Bv, bool need, int* map,
                                             const float* f, float k, float4* corr) {
    const unsigned long long bal = __ballot(need);
    if (!bal) return 0;   /* wave-uniform */
    const int m = __popcll(bal);
    const unsigned lane = gd_lane();
    if (need) map[__popcll(bal & gd_lanemask_lt())] = (int)lane;
    __builtin_amdgcn_wave_barrier();
    float cx, cy, cz;
    gd_mlp_cells(W, Bv, m, map, f, cx, cy, cz);
    const int src = (int)lane < m ? map[lane] : 0;
    const float ks = __shfl(k, src);
    if ((int)lane < m) corr[lane] = make_float4(ks * cx, ks * cy, ks * cz, __int_as_float(src));
    __builtin_amdgcn_wave_barrier();
    return m;
}

/* The learned walk's cell measure s (its features, its correction test
 * s / r > minRatio and the training samples'): the mass radius under the
 * mass-radius opening rule (wrad given), else the cell size ext / 2^level. */
__device__ __forceinline__ float gd_cell_measure(const float* __restrict__ wrad, float ext,
                                                 const int2* __restrict__ wlink, int idx) {
    return wrad ? wrad[idx] : ldexpf(ext, -(wlink[idx].y & 31));
}

/*
 * The network's inputs for a cell seen from point x: with d = c - x, r = |d|,
 * u = d / r, s the cell's edge, m its mass, O its third moments about c,
 * (O:uu)_i = O_ijk u_j u_k and T_i = O_ijj:
 *   f = [u (3), s / r, (7.5 O:uu - 1.5 T) / (m s^3) (3), u.O:uu / (m s^3)].
 * The cell's octupole term is (m s^3 / r^5) (v + u (20 a - 5 u.v)) with
 * v = f[4..6], a = f[7], so the leading part of what the quadrupole misses
 * is a linear function of the inputs; the network learns it and the
 * higher orders that correlate with it.  o0/o1/o2 hold O (xxx yyy zzz xxy,
 * xxz xyy yyz xzz, yzz xyz - -).  k = m s^3, or 0 when the cell is not
 * corrected (invalid, massless, or s / r <= minRatio).
 */
__device__ __forceinline__ void gd_cell_features(float4 c, float4 o0, float4 o1, float4 o2,
                                                 float size, float x, float y, float z,
                                                 bool valid, float minRatio, float* f,
                                                 float& k) {
    k = c.w * size * size * size;
    const float dx = c.x - x, dy = c.y - y, dz = c.z - z;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    const float rinv = valid ? __builtin_amdgcn_rsqf(r2) : 0.f;
    const float ux = dx * rinv, uy = dy * rinv, uz = dz * rinv;
    const float ratio = size * rinv;
    const bool use = valid && k > 0.f;
    const float k3inv = use ? __builtin_amdgcn_rcpf(k) : 0.f;
    const float xx = ux * ux, yy = uy * uy, zz = uz * uz;
    const float xy2 = 2.f * ux * uy, xz2 = 2.f * ux * uz, yz2 = 2.f * uy * uz;
    /* O_ijk u_j u_k and the trace vector T_i = O_ijj */
    const float ox = o0.x * xx + o1.y * yy + o1.w * zz + o0.w * xy2 + o1.x * xz2 + o2.y * yz2;
    const float oy = o0.w * xx + o0.y * yy + o2.x * zz + o1.y * xy2 + o2.y * xz2 + o1.z * yz2;
    const float oz = o1.x * xx + o1.z * yy + o0.z * zz + o2.y * xy2 + o1.w * xz2 + o2.x * yz2;
    const float tx = o0.x + o1.y + o1.w, ty = o0.w + o0.y + o2.x, tz = o1.x + o1.z + o0.z;
    f[0] = ux; f[1] = uy; f[2] = uz; f[3] = ratio;
    f[4] = (7.5f * ox - 1.5f * tx) * k3inv;
    f[5] = (7.5f * oy - 1.5f * ty) * k3inv;
    f[6] = (7.5f * oz - 1.5f * tz) * k3inv;
    f[7] = (ux * ox + uy * oy + uz * oz) * k3inv;
    if (!(use && ratio > minRatio)) k = 0.f;
}

/*
 * Evaluation of one group's list by one wavefront, in chunks of 64 entries.
 * Each lane gathers one entry's record into LDS (double buffered: the next
 * chunk's gather is in flight while the current one is evaluated), then all
 * 64 bodies evaluate the chunk with broadcast LDS reads; each chunk's partial
 * sum is added with Kahan compensation.  MODE 0: every entry is a monopole;
 * 1: back entries are quadrupole cells; 2: back entries are quadrupole cells
 * plus the network's residual correction (evaluated per chunk at the group's
 * centre on the matrix cores, scaled by each body's own m s^3 / r^5).
 * Only the corrected cells (s / r > minRatio, ~5 % of the cell entries at
 * theta 0.7) reach the per-body loop of the learned term, compacted into the
 * second half of sE (unused by the cell chunks), each pointing at its
 * record in the first half.  sE holds 2 x 64 float4; sQ0, sQ1 64 each;
 * sMap 64 ints.
 */
template <int MODE, bool PHI>
__device__ __forceinline__ void gd_eval_group(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wmom,
    const gd_tree_info* __restrict__ info, const int* __restrict__ list, int cap, int2 cnt,
    long long start, int count, float eps2, const gd_bf16x4* __restrict__ mlpW,
    const float* __restrict__ mlpB, float minRatio, float4* sE, float4* sQ0, float4* sQ1,
    int* sMap, gd_kahan3& tot, float& phs, float& phc,
    unsigned long long* nnEvals) {
    const unsigned lane = gd_lane();
    const bool valid = (int)lane < count;
    const float4 p = pos[start + (valid ? lane : 0)];
    const float px = p.x, py = p.y, pz = p.z;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    /* padding of a partial chunk: massless and far away, contributes exactly 0 */
    const float4 pad = make_float4(1e18f, 1e18f, 1e18f, 0.f);

    /* front list: monopoles, double-buffered gathers */
    const int nf = cnt.x;
    float4 nxt = pad;
    if ((int)lane < nf) nxt = wcom[list[lane]];
    int buf = 0;
    for (int c0 = 0; c0 < nf; c0 += 64) {
        sE[64 * buf + lane] = nxt;
        __builtin_amdgcn_wave_barrier();
        nxt = pad;
        if (c0 + 64 + (int)lane < nf) nxt = wcom[list[c0 + 64 + lane]];
        float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
        gd_eval64<PHI>(sE + 64 * buf, px, py, pz, eps2, ax, ay, az, ph);
        gd_kahan_add(tot.x, tot.cx, ax);
        gd_kahan_add(tot.y, tot.cy, ay);
        gd_kahan_add(tot.z, tot.cz, az);
        if (PHI) gd_kahan_add(phs, phc, ph);
        buf ^= 1;
    }
    if (MODE == 0) return;
    /* back list: cells */
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
        float4 c = pad, q0 = zero4, q1 = zero4;
        int idx = -1;
        if (e < nb) {
            idx = list[cap - 1 - e];
            c = wcom[idx];
            gd_mom_quad(wmom, idx, q0, q1);
        }
        float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
        sE[lane] = c;
        sQ0[lane] = make_float4(7.5f * q0.x, 7.5f * q0.y, 7.5f * q0.z, 7.5f * q0.w);
        sQ1[lane] = make_float4(7.5f * q1.x, 7.5f * q1.y, -1.5f * q1.z, q1.z);
        int nc = 0;
        if (MODE == 2) {
            /* network at the group centre, on the cells with s / r > minRatio */
            float4 o0 = zero4, o1 = zero4, o2 = zero4;
            float size = 1.f;
            if (idx >= 0) {
                gd_mom_oct(wmom, idx, o0, o1, o2);
                size = gd_cell_measure(wrad, ext, wlink, idx);
            }
            float f[8], k;
            gd_cell_features(c, o0, o1, o2, size, gx, gy, gz, idx >= 0, minRatio, f, k);
            nc = gd_mlp_compact(mlpW, mlpB, k > 0.f, sMap, f, k, sE + 64);
            *nnEvals += (unsigned long long)nc;
        }
        __builtin_amdgcn_wave_barrier();
        gd_eval64_quad<PHI>(sE, sQ0, sQ1, px, py, pz, eps2, ax, ay, az, ph);
        if (MODE == 2) {
            /* learned residual k / r^5 * o of each corrected cell */
            for (int kk = 0; kk < nc; ++kk) {
                const float4 oo = sE[64 + kk];
                const float4 cc = sE[__float_as_int(oo.w)];
                const float dx = cc.x - px, dy = cc.y - py, dz = cc.z - pz;
                const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                const float i1 = __builtin_amdgcn_rsqf(r2);
                const float i2 = i1 * i1;
                const float s5 = i2 * i2 * i1;
                ax = __builtin_fmaf(s5, oo.x, ax);
                ay = __builtin_fmaf(s5, oo.y, ay);
                az = __builtin_fmaf(s5, oo.z, az);
            }
        }
        gd_kahan_add(tot.x, tot.cx, ax);
        gd_kahan_add(tot.y, tot.cy, ay);
        gd_kahan_add(tot.z, tot.cz, az);
        if (PHI) gd_kahan_add(phs, phc, ph);
        __builtin_amdgcn_wave_barrier();
    }
}

/* Evaluation kernel: slot s (one wavefront) evaluates the list of group
 * groupList[s] (or s) when it has one (counts[s].x >= 0). */
template <int MODE, bool PHI>
__global__ __launch_bounds__(64 * GD_EV_WAVES) void k_walk_eval(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wmom,
    const gd_tree_info* __restrict__ info, const int2* __restrict__ slotRange,
    const int* __restrict__ groupList, int slots,
    const int* __restrict__ lists, const long long* __restrict__ offsets,
    const int* __restrict__ caps, const int2* __restrict__ counts, int cap,
    float3* __restrict__ acc, float* __restrict__ phi, float eps2, float G, int n,
    const gd_bf16x4* __restrict__ mlpW, const float* __restrict__ mlpB, float minRatio,
    unsigned long long* __restrict__ stats) {
    __shared__ float4 sE[GD_EV_WAVES][2 * 64];
    __shared__ float4 sQ0[MODE != 0 ? GD_EV_WAVES : 1][64];
    __shared__ float4 sQ1[MODE != 0 ? GD_EV_WAVES : 1][64];
    __shared__ int sMap[MODE == 2 ? GD_EV_WAVES : 1][64];
    __shared__ gd_bf16x4 sW[MODE == 2 ? GD_MLP_FWD_TILES * 64 : 1];
    __shared__ float sB[MODE == 2 ? GD_MLP_BIAS_PAD : 1];
    if (MODE == 2) {
        /* the network's packed weights and biases, once per block */
        for (int e = threadIdx.x; e < GD_MLP_FWD_TILES * 64; e += 64 * GD_EV_WAVES)
            sW[e] = mlpW[e];
        for (int e = threadIdx.x; e < GD_MLP_BIAS_PAD; e += 64 * GD_EV_WAVES) sB[e] = mlpB[e];
        __syncthreads();
        mlpW = sW;
        mlpB = sB;
    }
    const unsigned lane = gd_lane();
    const int wave = (int)(threadIdx.x >> 6);
    const long long s = (long long)gd_xcd_block() * GD_EV_WAVES + wave;
    if (s >= slots) return;
    const int2 cnt = counts[s];
    if (cnt.x < 0) return;   /* left to the next pass */
    const long long g = groupList ? groupList[s] : s;
    const int2 rg = slotRange[g];
    if (rg.y == 0) return;   /* empty slot */
    gd_kahan3 tot = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float phs = 0.f, phc = 0.f;
    unsigned long long nn = 0;
    const int* list = offsets ? lists + offsets[g] : lists + (size_t)s * cap;
    const int size = offsets ? caps[g] : cap;
    gd_eval_group<MODE, PHI>(pos, wcom, wlink, wrad, wmom, info, list, size, cnt, rg.x, rg.y,
                             eps2,
                             mlpW, mlpB, minRatio, sE[wave],
                             sQ0[MODE != 0 ? wave : 0], sQ1[MODE != 0 ? wave : 0],
                             sMap[MODE == 2 ? wave : 0], tot, phs,
                             phc, &nn);
    if (MODE == 2 && stats && lane == 0 && nn) atomicAdd(&gd_stat_slot(stats)[4], nn);
    if ((int)lane < rg.y) {
        const long long i = rg.x + lane;
        acc[i] = make_float3(G * tot.x, G * tot.y, G * tot.z);
        if (PHI) phi[i] = -G * phs;
    }
}

/*
 * Second-pass evaluation cut into segments.  The few groups that do not fit
 * the first pass have the longest lists; one wavefront each, they would be
 * evaluated nearly alone at the end of the walk.  Here slot s's list (of
 * lists[s * cap ..]) is cut into `segs` parts -- front (monopole) and back
 * (cell) entries alike -- and wavefront q evaluates part q % segs of slot
 * q / segs for the group's bodies, writing its per-body sums (x, y, z,
 * potential) to part[q * 64 + lane]; k_walk_eval_sum adds the parts in order.
 */
template <int MODE, bool PHI>
__global__ __launch_bounds__(64 * GD_EV_WAVES) void k_walk_eval_seg(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wmom,
    const gd_tree_info* __restrict__ info, const int2* __restrict__ slotRange,
    const int* __restrict__ groupList, int slots, int segs,
    const int* __restrict__ lists, const int2* __restrict__ counts, int cap,
    float4* __restrict__ part, float eps2,
    const gd_bf16x4* __restrict__ mlpW, const float* __restrict__ mlpB, float minRatio,
    unsigned long long* __restrict__ stats, const int* __restrict__ slotsDev = nullptr) {
    __shared__ float4 sE[GD_EV_WAVES][2 * 64];
    __shared__ float4 sQ0[MODE != 0 ? GD_EV_WAVES : 1][64];
    __shared__ float4 sQ1[MODE != 0 ? GD_EV_WAVES : 1][64];
    __shared__ int sMap[MODE == 2 ? GD_EV_WAVES : 1][64];
    __shared__ gd_bf16x4 sW[MODE == 2 ? GD_MLP_FWD_TILES * 64 : 1];
    __shared__ float sB[MODE == 2 ? GD_MLP_BIAS_PAD : 1];
    if (MODE == 2) {
        for (int e = threadIdx.x; e < GD_MLP_FWD_TILES * 64; e += 64 * GD_EV_WAVES)
            sW[e] = mlpW[e];
        for (int e = threadIdx.x; e < GD_MLP_BIAS_PAD; e += 64 * GD_EV_WAVES) sB[e] = mlpB[e];
        __syncthreads();
        mlpW = sW;
        mlpB = sB;
    }
    const unsigned lane = gd_lane();
    const int wave = (int)(threadIdx.x >> 6);
    const long long q = (long long)gd_xcd_block() * GD_EV_WAVES + wave;
    const long long used = gd_slots(slots, slotsDev);
    if (segs <= 0) segs = gd_seg_count((int)used);   /* from the device-held count */
    if (q >= used * segs) return;
    const long long s = q / segs;
    const int seg = (int)(q - s * segs);
    const int2 cnt = counts[s];
    if (cnt.x < 0) return;   /* left to the per-body walk */
    const int2 rg = slotRange[groupList[s]];
    if (rg.y == 0) return;
    const int f0 = (int)((long long)cnt.x * seg / segs);
    const int f1 = (int)((long long)cnt.x * (seg + 1) / segs);
    const int b0 = (int)((long long)cnt.y * seg / segs);
    const int b1 = (int)((long long)cnt.y * (seg + 1) / segs);
    /* front entries f0.. at list[f0..]; back entry e of the part is the
       list's back entry b0 + e, at lists[s * cap + cap - 1 - b0 - e] */
    gd_kahan3 tot = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float phs = 0.f, phc = 0.f;
    unsigned long long nn = 0;
    gd_eval_group<MODE, PHI>(pos, wcom, wlink, wrad, wmom, info,
                             lists + (size_t)s * cap + f0, cap - f0 - b0,
                             make_int2(f1 - f0, b1 - b0), rg.x, rg.y, eps2, mlpW, mlpB,
                             minRatio, sE[wave], sQ0[MODE != 0 ? wave : 0],
                             sQ1[MODE != 0 ? wave : 0], sMap[MODE == 2 ? wave : 0], tot, phs,
                             phc, &nn);
    if (MODE == 2 && stats && lane == 0 && nn) atomicAdd(&gd_stat_slot(stats)[4], nn);
    part[q * 64 + lane] = make_float4(tot.x, tot.y, tot.z, phs);
}

/* The segments' sums of each second-pass group, added in segment order with
 * compensation: one wavefront per slot, lane = body. */
template <bool PHI>
__global__ __launch_bounds__(256) void k_walk_eval_sum(const int2* __restrict__ slotRange,
                                                       const int* __restrict__ groupList,
                                                       int slots, int segs,
                                                       const int2* __restrict__ counts,
                                                       const float4* __restrict__ part,
                                                       float3* __restrict__ acc,
                                                       float* __restrict__ phi, float G,
                                                       const int* __restrict__ slotsDev =
                                                           nullptr) {
    const unsigned lane = gd_lane();
    const long long s = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long long used = gd_slots(slots, slotsDev);
    if (segs <= 0) segs = gd_seg_count((int)used);
    if (s >= used) return;
    if (counts[s].x < 0) return;
    const int2 rg = slotRange[groupList[s]];
    if ((int)lane >= rg.y) return;
    gd_kahan3 tot = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float phs = 0.f, phc = 0.f;
#pragma unroll 8
    for (int k = 0; k < segs; ++k) {
        const float4 v = part[((size_t)s * segs + k) * 64 + lane];
        gd_kahan_add(tot.x, tot.cx, v.x);
        gd_kahan_add(tot.y, tot.cy, v.y);
        gd_kahan_add(tot.z, tot.cz, v.z);
        if (PHI) gd_kahan_add(phs, phc, v.w);
    }
    const long long i = rg.x + lane;
    acc[i] = make_float3(G * tot.x, G * tot.y, G * tot.z);
    if (PHI) phi[i] = -G * phs;
}

/*
 * Learned corrections of a wavefront's queued cells (first pass, learned
 * mode): queue[0 .. qn) holds node indices, bit 30 set for the cells of half
 * B.  Lane l < qn computes cell l's features at its half's centre, the
 * network runs on the qn cells at once (full 16-row tiles instead of a few
 * cells per chunk), and every body adds k / r^5 * output of each cell of its
 * own half (k = m s^3, r from the body) to (ax, ay, az), reading its half's
 * records (centre, k * output) from rec, so both halves' lanes work in the
 * same steps.  map: 64 ints, rec: 256 float4 of LDS.  Returns the number of
 * network evaluations.
 */
__device__ __forceinline__ int gd_flush_corr(
    const float4* __restrict__ wcom, const int2* __restrict__ wlink,
    const float* __restrict__ wrad,
    const float4* __restrict__ wmom, float ext, const gd_bf16x4* __restrict__ W,
    const float* __restrict__ Bv, float minRatio, const int* queue, int qn, int* map,
    float gax, float gay, float gaz, float gbx, float gby, float gbz, bool inB, float px,
    float py, float pz, float4* rec, float& ax, float& ay, float& az) {
    const unsigned lane = gd_lane();
    const bool have = (int)lane < qn;
    const int e = have ? queue[lane] : 0;
    const int idx = e & 0x3fffffff;
    const bool half = (e >> 30) & 1;
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f), o0 = c, o1 = c, o2 = c;
    float size = 1.f;
    if (have) {
        c = wcom[idx];
        gd_mom_oct(wmom, idx, o0, o1, o2);
        size = gd_cell_measure(wrad, ext, wlink, idx);
    }
    float f[8], k;
    gd_cell_features(c, o0, o1, o2, size, half ? gbx : gax, half ? gby : gay,
                     half ? gbz : gaz, have, minRatio, f, k);
    if (have) map[lane] = (int)lane;
    __builtin_amdgcn_wave_barrier();
    float cx, cy, cz;
    gd_mlp_cells(W, Bv, qn, map, f, cx, cy, cz);
    /* each half's corrected cells as records (centre, k * output) in LDS, the
       shorter half padded with far-away zero entries; every lane then reads
       its own half's records, both halves in the same steps */
    const unsigned long long lt = gd_lanemask_lt();
    const unsigned long long balA = __ballot(have && !half), balB = __ballot(have && half);
    const int nA = __popcll(balA), nB = __popcll(balB);
    const int nmax = nA > nB ? nA : nB;
    if (have) {
        const int j = __popcll((half ? balB : balA) & lt);
        float4* r = rec + (half ? 128 : 0) + 2 * j;
        r[0] = make_float4(c.x, c.y, c.z, 0.f);
        r[1] = make_float4(k * cx, k * cy, k * cz, 0.f);
    }
    const float4 far = make_float4(1e18f, 1e18f, 1e18f, 0.f), zero = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((int)lane >= nA && (int)lane < nmax) {
        rec[2 * lane] = far;
        rec[2 * lane + 1] = zero;
    }
    if ((int)lane >= nB && (int)lane < nmax) {
        rec[128 + 2 * lane] = far;
        rec[128 + 2 * lane + 1] = zero;
    }
    __builtin_amdgcn_wave_barrier();
    const float4* my = rec + (inB ? 128 : 0);
    for (int kk = 0; kk < nmax; ++kk) {
        const float4 q = my[2 * kk], o = my[2 * kk + 1];
        const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const float i1 = __builtin_amdgcn_rsqf(r2);
        const float i2 = i1 * i1;
        const float s5 = i2 * i2 * i1;
        ax = __builtin_fmaf(s5, o.x, ax);
        ay = __builtin_fmaf(s5, o.y, ay);
        az = __builtin_fmaf(s5, o.z, az);
    }
    __builtin_amdgcn_wave_barrier();
    return qn;
}

/*
 * The cells of a half's list in the far split (entries 0..n-1 at
 * list[cap - 1 - e]).  The 64 lanes take 64 entries at a time (indices two
 * chunks ahead, records one chunk ahead, so the gathers overlap the work;
 * two chunks per iteration in alternating registers); a cell whose centre of
 * mass is far from the half's box (h2 < far2 R^2, R from the box centre g)
 * goes with its record to the LDS ring sRec (128 entries of 3 float4: fq
 * pending from entry fh on, < 64 between calls), and each time 64 are
 * pending every lane adds one of them to its local-expansion coefficients L
 * (gd_far.h); the near cells are written back, in order, to the front of the
 * region, for the evaluation's per-body loop.  KEEP (learned walk): a cell
 * the network corrects (s / r > minRatio from g, the evaluation's own test)
 * stays near, so the evaluation finds it.  Returns the near count.
 */
#define GD_FAR_RING 128
template <bool KEEP>
__device__ __forceinline__ int gd_far_region(const float4* __restrict__ wcom,
                                             const int2* __restrict__ wlink,
                                             const float* __restrict__ wrad,
                                             const float4* __restrict__ wmom, float ext,
                                             int* __restrict__ list, int cap, int n, float gx,
                                             float gy, float gz, float h2, float far2, float eps2,
                                             float minRatio, float4* sRec, gd_local& L, int& fq,
                                             int& fh, int& nFar) {
    const unsigned lane = gd_lane();
    const unsigned long long lt = gd_lanemask_lt();
    /* entry e of the region */
    auto at = [&](int e) -> int& { return list[cap - 1 - e]; };
    /* the entry's index (-1 past the end); a lane past the end gathers node
       0's records (never used), so the gathers need no branch */
    auto index = [&](int e) -> int { return e < n ? at(e) : -1; };
    auto gather = [&](int idx, float4& c, float4& q0, float4& q1) {
        const int safe = idx >= 0 ? idx : 0;
        c = wcom[safe];
        gd_mom_quad(wmom, safe, q0, q1);
    };
    int nNear = 0;
    /* one chunk: far test, near write-back, far records to the ring, and a
       wave of expansion adds when 64 are pending */
    auto chunk = [&](int idx, float4 c, float4 q0, float4 q1) {
        const float dx = c.x - gx, dy = c.y - gy, dz = c.z - gz;
        const float r2 = (dx * dx + dy * dy) + dz * dz;
        bool isFar = idx >= 0 && h2 < far2 * r2;
        if (KEEP && isFar && c.w > 0.f) {
            const float rq = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
            const float size = gd_cell_measure(wrad, ext, wlink, idx);
            if (size * __builtin_amdgcn_rsqf(rq) > minRatio) isFar = false;
        }
        const bool isNear = idx >= 0 && !isFar;
        const unsigned long long nbal = __ballot(isNear);
        /* in place: entry e moves to nNear + rank <= e, already read by this wave */
        if (isNear) at(nNear + __popcll(nbal & lt)) = idx;
        nNear += __popcll(nbal);
        const unsigned long long fbal = __ballot(isFar);
        if (isFar) {
            float4* q = sRec + 3 * ((fh + fq + __popcll(fbal & lt)) & (GD_FAR_RING - 1));
            q[0] = c;
            q[1] = q0;
            q[2] = q1;
        }
        fq += __popcll(fbal);
        __builtin_amdgcn_wave_barrier();
        if (fq >= 64) {   /* wave-uniform: a full wave of far cells */
            const float4* q = sRec + 3 * ((fh + (int)lane) & (GD_FAR_RING - 1));
            const float4 cj = q[0], a = q[1], b = q[2];
            gd_local_add(L, gx - cj.x, gy - cj.y, gz - cj.z, cj.w, a, b, eps2);
            nFar += 64;
            fh = (fh + 64) & (GD_FAR_RING - 1);
            fq -= 64;
            __builtin_amdgcn_wave_barrier();
        }
    };
    /* chunk 0's records, chunk 1's indices */
    int iA = index((int)lane), iB = index(64 + (int)lane);
    float4 cA, qA0, qA1, cB, qB0, qB1;
    gather(iA, cA, qA0, qA1);
    for (int c0 = 0; c0 < n; c0 += 128) {
        /* indices of chunk c0 + 128 and records of chunk c0 + 64 (entries >=
           c0 + 64: the write-backs of chunk c0 go to entries < c0 + 64) */
        const int iA2 = index(c0 + 128 + (int)lane);
        gather(iB, cB, qB0, qB1);
        chunk(iA, cA, qA0, qA1);
        if (c0 + 64 >= n) break;   /* wave-uniform */
        const int iB2 = index(c0 + 192 + (int)lane);
        gather(iA2, cA, qA0, qA1);
        chunk(iB, cB, qB0, qB1);
        iA = iA2;
        iB = iB2;
    }
    return nNear;
}

/*
 * Far split of one half's list by the run's wavefront: its cells (nb, at the
 * back) through gd_far_region; then the queue's last entries, the sum of the
 * lanes' coefficients, and the expansion at each body's position, added to
 * (fx, fy, fz) by the lanes with `mine` (the half's bodies).  sRec: the queue
 * (128 x 3 float4), then the reduction (64 x GD_LOCAL_STRIDE floats).  Writes
 * the near counts to nearF (the leaves, nf: all near) and nearB.
 */
template <bool KEEP>
__device__ __forceinline__ void gd_far_half(const float4* __restrict__ wcom,
                                            const int2* __restrict__ wlink,
                                            const float* __restrict__ wrad,
                                            const float4* __restrict__ wmom, float ext,
                                            int* __restrict__ list, int cap, int nf, int nb,
                                            float gx, float gy, float gz, float h2, float far2,
                                            float eps2, float minRatio, bool mine, float px,
                                            float py, float pz, float4* sRec, int& nearF,
                                            int& nearB, float& fx, float& fy, float& fz) {
    const unsigned lane = gd_lane();
    gd_local L;
    gd_local_zero(L);
    int fq = 0, fh = 0, nFar = 0;
    /* leaves stay in the per-body loop: they are mostly near, and streaming
       them through the split cost more than the far ones saved */
    nearF = nf;
    nearB = gd_far_region<KEEP>(wcom, wlink, wrad, wmom, ext, list, cap, nb, gx, gy, gz, h2, far2,
                                eps2, minRatio, sRec, L, fq, fh, nFar);
    if (fq > 0) {
        if ((int)lane < fq) {
            const float4* q = sRec + 3 * ((fh + (int)lane) & (GD_FAR_RING - 1));
            const float4 cj = q[0], a = q[1], b = q[2];
            gd_local_add(L, gx - cj.x, gy - cj.y, gz - cj.z, cj.w, a, b, eps2);
        }
        nFar += fq;
        __builtin_amdgcn_wave_barrier();
    }
    if (nFar > 0) {   /* wave-uniform */
        /* the queue's records are consumed: its LDS holds the reduction */
        gd_local_allreduce(L, reinterpret_cast<float*>(sRec));
        gd_local_finish(L);
        float ax = 0.f, ay = 0.f, az = 0.f;
        gd_local_eval(L, px - gx, py - gy, pz - gz, ax, ay, az);
        if (mine) {
            fx += ax;
            fy += ay;
            fz += az;
        }
    }
}

/*
 * Far-field kernel of the first pass (runs before k_walk_eval_pair, with the
 * same run-to-wavefront mapping): for each half of run r with a list, the far
 * cells are summed through the half's local expansion (gd_far_half), the
 * half's cell count becomes its near count (the near cells compacted in
 * place), and each body's far field (without G) is written to acc, which the
 * evaluation then adds to.  KEEP: learned walk.
 */
template <int EVW, bool KEEP>
__global__ __launch_bounds__(64 * EVW) void k_walk_far(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wmom,
    const gd_tree_info* __restrict__ info, const int2* __restrict__ slotRange, int runs,
    int* __restrict__ lists, const long long* __restrict__ offsets,
    const int* __restrict__ caps, int2* __restrict__ counts, float3* __restrict__ acc,
    float eps2, float far2, float minRatio, unsigned long long* __restrict__ stats) {
    /* per wavefront: the far queue (128 x 3 float4), later the reduction */
    constexpr int REC = (3 * 128 > 16 * GD_LOCAL_STRIDE) ? 3 * 128 : 16 * GD_LOCAL_STRIDE;
    __shared__ float4 sRec[EVW][REC];
    if (info->overflow) return;   /* block-uniform: no tree, nothing listed */
    const unsigned lane = gd_lane();
    const int wave = (int)(threadIdx.x >> 6);
    const long long r = (long long)gd_xcd_block() * EVW + wave;
    if (r >= runs) return;   /* wave-uniform */
    const int2 rgA = slotRange[2 * r], rgB = slotRange[2 * r + 1];
    const int2 cA = counts[2 * r];
    const int2 cB = rgB.y > 0 ? counts[2 * r + 1] : make_int2(-1, -1);
    const bool actA = cA.x >= 0, actB = cB.x >= 0;
    if (!actA && !actB) return;
    const int nA = rgA.y, nB = rgB.y;
    const bool valid = (int)lane < nA + nB;
    const bool inB = (int)lane >= nA;
    const float4 p = pos[rgA.x + (valid ? lane : 0)];
    const float ext = info->box.extent;
    float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        const bool act = h ? actB : actA;
        const int2 cn = h ? cB : cA;
        if (!act || cn.y <= 0) continue;   /* wave-uniform */
        float lx, ly, lz, hx, hy, hz;
        gd_masked_box(valid && (h ? inB : !inB), p.x, p.y, p.z, lx, ly, lz, hx, hy, hz);
        const float gx = 0.5f * (lx + hx), gy = 0.5f * (ly + hy), gz = 0.5f * (lz + hz);
        const float ex = hx - lx, ey = hy - ly, ez = hz - lz;
        const float h2 = 0.25f * ((ex * ex + ey * ey) + ez * ez);
        const long long s = 2 * r + h;
        int nearF, nearB;
        gd_far_half<KEEP>(wcom, wlink, wrad, wmom, ext, lists + offsets[s], caps[s], cn.x, cn.y, gx, gy,
                          gz, h2, far2, eps2, minRatio, valid && (h ? inB : !inB), p.x, p.y, p.z,
                          sRec[wave], nearF, nearB, fx, fy, fz);
        if (lane == 0) {
            counts[s] = make_int2(nearF, nearB);
            const int moved = cn.x + cn.y - nearF - nearB;
            if (stats && moved > 0) {
                unsigned long long* slot = gd_stat_slot(stats);
                atomicAdd(&slot[5], (unsigned long long)moved);
                atomicAdd(&slot[6], (unsigned long long)moved * (unsigned long long)(h ? nB : nA));
            }
        }
    }
    if (valid && (inB ? actB : actA)) acc[rgA.x + lane] = make_float3(fx, fy, fz);
}

/*
 * First-pass evaluation of one 64-body run by one wavefront: the run's two
 * group slots (halves A and B; B may be empty) each have their own list.
 * Lane l is body start + l and belongs to A (l < nA) or B; per chunk step
 * each lane gathers entry c0 + l of both lists into LDS (sA, sB), and every
 * lane evaluates the 64 entries of its own half's chunk.  The shorter list
 * is padded (massless, far away: exactly 0), so the run costs the longer of
 * the two lists instead of their sum.  cA / cB are the halves' counts, (0, 0)
 * for a half that is not evaluated here (empty, or left to the later passes).
 * In learned mode the cells to correct (s / r > minRatio from their half's
 * centre) are queued (sQueue, 64 ints) and the network runs on full batches
 * (gd_flush_corr).  sQ holds 384 float4: a cell chunk's records, three
 * float4 per entry and half (sA, sB alias its start for the monopole chunks),
 * and between chunks the flush's per-half correction records.
 */
template <int MODE, bool PHI>
__device__ __forceinline__ void gd_eval_pair(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wmom, const gd_tree_info* __restrict__ info,
    const int* __restrict__ listA, int capA, int2 cA, const int* __restrict__ listB, int capB,
    int2 cB, long long start, int nA, int nB, float eps2, const gd_bf16x4* __restrict__ mlpW,
    const float* __restrict__ mlpB, float minRatio, float4* sA, float4* sB, float4* sQ,
    int* sMap, int* sQueue, gd_kahan3& tot, float& phs, float& phc,
    unsigned long long* nnEvals) {
    const unsigned lane = gd_lane();
    const bool valid = (int)lane < nA + nB;
    const bool inB = (int)lane >= nA;
    const float4 p = pos[start + (valid ? lane : 0)];
    const float px = p.x, py = p.y, pz = p.z;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 pad = make_float4(1e18f, 1e18f, 1e18f, 0.f);
    const float4* myE = inB ? sB : sA;

    /* front lists: monopoles */
    const int nfA = cA.x, nfB = cB.x;
    const int nf = nfA > nfB ? nfA : nfB;
    float4 na = pad, nb = pad;
    if ((int)lane < nfA) na = wcom[listA[lane]];
    if ((int)lane < nfB) nb = wcom[listB[lane]];
    for (int c0 = 0; c0 < nf; c0 += 64) {
        sA[lane] = na;
        sB[lane] = nb;
        __builtin_amdgcn_wave_barrier();
        na = pad;
        nb = pad;
        if (c0 + 64 + (int)lane < nfA) na = wcom[listA[c0 + 64 + lane]];
        if (c0 + 64 + (int)lane < nfB) nb = wcom[listB[c0 + 64 + lane]];
        float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
        gd_eval64<PHI>(myE, px, py, pz, eps2, ax, ay, az, ph);
        gd_kahan_add(tot.x, tot.cx, ax);
        gd_kahan_add(tot.y, tot.cy, ay);
        gd_kahan_add(tot.z, tot.cz, az);
        if (PHI) gd_kahan_add(phs, phc, ph);
        __builtin_amdgcn_wave_barrier();
    }
    if (MODE == 0) return;

    /* back lists: cells */
    const int nbA = cA.y, nbB = cB.y;
    const int nbm = nbA > nbB ? nbA : nbB;
    float gax = 0.f, gay = 0.f, gaz = 0.f, gbx = 0.f, gby = 0.f, gbz = 0.f;
    if (MODE == 2) {
        float lx, ly, lz, hx, hy, hz;
        gd_masked_box(valid && !inB, px, py, pz, lx, ly, lz, hx, hy, hz);
        if (nA > 0) { gax = 0.5f * (lx + hx); gay = 0.5f * (ly + hy); gaz = 0.5f * (lz + hz); }
        gd_masked_box(valid && inB, px, py, pz, lx, ly, lz, hx, hy, hz);
        if (nB > 0) { gbx = 0.5f * (lx + hx); gby = 0.5f * (ly + hy); gbz = 0.5f * (lz + hz); }
    }
    const float ext = info->box.extent;
    /* cell records: half A's 64 at sQ[0 .. 191], half B's at sQ[192 ..], three
       float4 each (the monopole buffers sA, sB alias the start of sQ, done) */
    float4* recA = sQ + 3 * lane;
    float4* recB = sQ + 192 + 3 * lane;
    const float4* myRec = sQ + (inB ? 192 : 0);
    int qn = 0;   /* queued cells (learned mode) */
    float lx = 0.f, ly = 0.f, lz = 0.f;   /* their corrections, since the last Kahan add */
    for (int c0 = 0; c0 < nbm; c0 += 64) {
        const int e = c0 + (int)lane;
        float4 ca = pad, qa0 = zero4, qa1 = zero4, cb = pad, qb0 = zero4, qb1 = zero4;
        int ia = -1, ib = -1;
        if (e < nbA) {
            ia = listA[capA - 1 - e];
            ca = wcom[ia];
            gd_mom_quad(wmom, ia, qa0, qa1);
        }
        if (e < nbB) {
            ib = listB[capB - 1 - e];
            cb = wcom[ib];
            gd_mom_quad(wmom, ib, qb0, qb1);
        }
        recA[0] = ca;
        recA[1] = make_float4(7.5f * qa0.x, 7.5f * qa0.y, 7.5f * qa0.z, 7.5f * qa0.w);
        recA[2] = make_float4(7.5f * qa1.x, 7.5f * qa1.y, -1.5f * qa1.z, qa1.z);
        recB[0] = cb;
        recB[1] = make_float4(7.5f * qb0.x, 7.5f * qb0.y, 7.5f * qb0.z, 7.5f * qb0.w);
        recB[2] = make_float4(7.5f * qb1.x, 7.5f * qb1.y, -1.5f * qb1.z, qb1.z);
        __builtin_amdgcn_wave_barrier();
        float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
        gd_eval64_quad3<PHI>(myRec, px, py, pz, eps2, ax, ay, az, ph);
        if (MODE == 2) {
            /* queue the cells each half corrects: s / r > minRatio from its centre,
               with the feature function's own rounding of s / r */
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int id = h ? ib : ia;
                const float4 cc = h ? cb : ca;
                bool need = false;
                if (id >= 0 && cc.w > 0.f) {
                    const float dx = cc.x - (h ? gbx : gax), dy = cc.y - (h ? gby : gay),
                                dz = cc.z - (h ? gbz : gaz);
                    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                    const float size = gd_cell_measure(wrad, ext, wlink, id);
                    need = size * __builtin_amdgcn_rsqf(r2) > minRatio;
                }
                const unsigned long long bal = __ballot(need);
                const int m = __popcll(bal);
                if (qn + m > 64) {   /* wave-uniform */
                    *nnEvals += (unsigned long long)gd_flush_corr(
                        wcom, wlink, wrad, wmom, ext, mlpW, mlpB, minRatio, sQueue, qn, sMap, gax,
                        gay, gaz, gbx, gby, gbz, inB, px, py, pz, sQ, lx, ly, lz);
                    qn = 0;
                }
                if (need) sQueue[qn + __popcll(bal & gd_lanemask_lt())] = id | (h << 30);
                qn += m;
                __builtin_amdgcn_wave_barrier();
            }
        }
        gd_kahan_add(tot.x, tot.cx, ax);
        gd_kahan_add(tot.y, tot.cy, ay);
        gd_kahan_add(tot.z, tot.cz, az);
        if (PHI) gd_kahan_add(phs, phc, ph);
        __builtin_amdgcn_wave_barrier();
    }
    if (MODE == 2) {
        if (qn > 0)
            *nnEvals += (unsigned long long)gd_flush_corr(
                wcom, wlink, wrad, wmom, ext, mlpW, mlpB, minRatio, sQueue, qn, sMap, gax, gay, gaz,
                gbx, gby, gbz, inB, px, py, pz, sQ, lx, ly, lz);
        gd_kahan_add(tot.x, tot.cx, lx);
        gd_kahan_add(tot.y, tot.cy, ly);
        gd_kahan_add(tot.z, tot.cz, lz);
    }
}

/* First-pass evaluation kernel: one wavefront per 64-body run (its two group
 * slots, gd_eval_pair); EVW wavefronts per block.  A half whose traversal
 * did not list it (counts -1: second pass or per-body walk) is skipped and
 * its bodies are written by those passes. */
template <int MODE, bool PHI, int EVW, bool FAR>
__global__ __launch_bounds__(64 * EVW) void k_walk_eval_pair(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wmom, const gd_tree_info* __restrict__ info,
    const int2* __restrict__ slotRange, int runs, const int* __restrict__ lists,
    const long long* __restrict__ offsets, const int* __restrict__ caps,
    const int2* __restrict__ counts, float3* __restrict__ acc, float* __restrict__ phi,
    float eps2, float G, const gd_bf16x4* __restrict__ mlpW, const float* __restrict__ mlpB,
    float minRatio, unsigned long long* __restrict__ stats) {
    /* per wavefront: monopole chunks of both halves (2 x 64 float4), then,
       reusing the same memory, cell records (2 x 64 x 3 float4) */
    __shared__ float4 sE[EVW][MODE != 0 ? 384 : 128];
    __shared__ int sMap[MODE == 2 ? EVW : 1][64];
    __shared__ int sQueue[MODE == 2 ? EVW : 1][64];
    __shared__ gd_bf16x4 sW[MODE == 2 ? GD_MLP_FWD_TILES * 64 : 1];
    __shared__ float sBias[MODE == 2 ? GD_MLP_BIAS_PAD : 1];
    if (info->overflow) return;   /* block-uniform: no tree, nothing listed */
    if (MODE == 2) {
        /* the network's packed weights and biases, once per block */
        for (int e = threadIdx.x; e < GD_MLP_FWD_TILES * 64; e += 64 * EVW) sW[e] = mlpW[e];
        for (int e = threadIdx.x; e < GD_MLP_BIAS_PAD; e += 64 * EVW) sBias[e] = mlpB[e];
        __syncthreads();
        mlpW = sW;
        mlpB = sBias;
    }
    const unsigned lane = gd_lane();
    const int wave = (int)(threadIdx.x >> 6);
    const long long r = (long long)gd_xcd_block() * EVW + wave;
    if (r >= runs) return;   /* wave-uniform */
    const int2 rgA = slotRange[2 * r], rgB = slotRange[2 * r + 1];
    int2 cA = counts[2 * r];
    int2 cB = rgB.y > 0 ? counts[2 * r + 1] : make_int2(-1, -1);
    const bool actA = cA.x >= 0, actB = cB.x >= 0;
    if (!actA && !actB) return;
    if (!actA) cA = make_int2(0, 0);
    if (!actB) cB = make_int2(0, 0);
    const int* listA = lists + (actA ? offsets[2 * r] : 0);
    const int* listB = lists + (actB ? offsets[2 * r + 1] : 0);
    const int capA = actA ? caps[2 * r] : 0, capB = actB ? caps[2 * r + 1] : 0;
    gd_kahan3 tot = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float phs = 0.f, phc = 0.f;
    unsigned long long nn = 0;
    gd_eval_pair<MODE, PHI>(pos, wcom, wlink, wrad, wmom, info, listA, capA, cA, listB, capB,
                            cB, rgA.x, rgA.y, rgB.y, eps2, mlpW, mlpB, minRatio, sE[wave],
                            sE[wave] + 64, sE[wave], sMap[MODE == 2 ? wave : 0],
                            sQueue[MODE == 2 ? wave : 0], tot, phs, phc, &nn);
    if (MODE == 2 && stats && lane == 0 && nn) atomicAdd(&gd_stat_slot(stats)[4], nn);
    const bool inB = (int)lane >= rgA.y;
    if ((int)lane < rgA.y + rgB.y && (inB ? actB : actA)) {
        const long long i = rgA.x + lane;
        if (FAR) {
            /* the far cells' field, from k_walk_far */
            const float3 f = acc[i];
            gd_kahan_add(tot.x, tot.cx, f.x);
            gd_kahan_add(tot.y, tot.cy, f.y);
            gd_kahan_add(tot.z, tot.cz, f.z);
        }
        acc[i] = make_float3(G * tot.x, G * tot.y, G * tot.z);
        if (PHI) phi[i] = -G * phs;
    }
}

/*
 * Per-body walk of the groups the traversal could not list (a group whose
 * bodies are spread out -- typically a run of the Morton curve that jumps
 * between distant clusters -- sees a huge list through its bounding box).
 * One wavefront walks one body: 64 nodes are tested per step against the
 * body itself, each accepted node is evaluated at once by the lane that
 * tested it (monopole; in MODE 1 quadrupole for cells), per-lane sums are
 * Kahan-compensated and combined by a fixed butterfly at the end.  Slot q
 * walks body bodyList[q], or, without bodyList, body q % 64 of group (slot)
 * groupList[q / 64] (when the group has that many bodies); WAVES wavefronts per block, each with an LDS stack of
 * STACK entries.  A body whose stack would overflow goes to overflowBodies
 * (when given; else the host reports an error): with STACK = GD_BODY_STACK
 * that cannot happen, since a step pops up to 64 nodes and pushes at most 8
 * children each, and the pushed segments along the stack have strictly
 * increasing minimum level, so the stack never exceeds 22 x 512 entries.
 * gd_walk_one_body walks body i with one wavefront; k_walk_body hands
 * slots to wavefronts.
 */
template <int MODE, bool PHI, int STACK>
__device__ __forceinline__ void gd_walk_one_body(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wmom,
    const gd_tree_info* __restrict__ info, long long i, int* sStack,
    int* __restrict__ overflowBodies, int* __restrict__ overflowCount,
    float3* __restrict__ acc, float* __restrict__ phi, float eps2, float G, float theta,
    unsigned long long* __restrict__ stats) {
    const unsigned lane = gd_lane();
    const float4 p = pos[i];
    const float ext = info->box.extent;
    const float th2 = theta * theta;
    const unsigned long long lt = gd_lanemask_lt();
    gd_kahan3 tot = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float phs = 0.f, phc = 0.f;
    unsigned long long nInter = 0, nCell = 0, nOpen = 0;
    int top = 1;
    if (lane == 0) sStack[0] = 0;
    while (top > 0) {
        const int take = top < 64 ? top : 64;
        const int base = top - take;
        bool accept = false, open = false, leaf = false;
        int idx = 0, nk = 0, first = 0;
        float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
        float size = 0.f, d2 = 0.f;
        if ((int)lane < take) {
            idx = sStack[base + lane];
            c = wcom[idx];
            const int2 lk = wlink[idx];
            size = ldexpf(ext, -(lk.y & 31));
            const float b = wrad ? wrad[idx] : size;   /* the opening rule's measure */
            const float dx = c.x - p.x, dy = c.y - p.y, dz = c.z - p.z;
            d2 = dx * dx + dy * dy + dz * dz;
            leaf = lk.x < 0;
            accept = leaf || (b * b < th2 * d2);
            open = !accept;
            nk = open ? (lk.y >> 5) : 0;
            first = lk.x;
        }
        __builtin_amdgcn_wave_barrier();
        top = base;
        int excl = 0, total = 0;
#pragma unroll
        for (int bit = 0; bit < 4; ++bit) {
            const unsigned long long bal = __ballot((nk >> bit) & 1);
            excl += __popcll(bal & lt) << bit;
            total += __popcll(bal) << bit;
        }
        nOpen += (unsigned long long)__popcll(__ballot(open));
        if (top + total > STACK) {
            if (lane == 0) {
                if (overflowBodies) overflowBodies[atomicAdd(overflowCount, 1)] = (int)i;
                else atomicOr(&gd_stat_slot(stats)[2], 1ull);   /* reported by the host */
            }
            return;
        }
        if (open)
            for (int k = 0; k < nk; ++k) sStack[top + excl + k] = first + k;
        top += total;
        /* evaluate the accepted nodes, one per lane */
        const bool cell = MODE != 0 && accept && !leaf;
        float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
        const float dx = c.x - p.x, dy = c.y - p.y, dz = c.z - p.z;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float mi = c.w * inv;
        if (accept && !cell) {
            const float f = mi * (inv * inv);
            ax = f * dx; ay = f * dy; az = f * dz;
            ph = mi;
        }
        if (MODE != 0) {
            const unsigned long long cb = __ballot(cell);
            nCell += (unsigned long long)__popcll(cb);
            float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
            if (cell) {
                gd_mom_quad(wmom, idx, q0, q1);
                const float inv2 = inv * inv;
                const float inv5 = inv * inv2 * inv2;
                const float sdx = __builtin_fmaf(q0.x, dx, __builtin_fmaf(q0.w, dy, q1.x * dz));
                const float sdy = __builtin_fmaf(q0.w, dx, __builtin_fmaf(q0.y, dy, q1.y * dz));
                const float sdz = __builtin_fmaf(q1.x, dx, __builtin_fmaf(q1.y, dy, q0.z * dz));
                const float q = __builtin_fmaf(dx, sdx, __builtin_fmaf(dy, sdy, dz * sdz));
                const float cd = __builtin_fmaf(mi, inv2, inv5 * __builtin_fmaf(7.5f * q, inv2, -1.5f * q1.z));
                const float cs = -3.f * inv5;
                ax = __builtin_fmaf(cd, dx, cs * sdx);
                ay = __builtin_fmaf(cd, dy, cs * sdy);
                az = __builtin_fmaf(cd, dz, cs * sdz);
                ph = __builtin_fmaf(0.5f * inv5, __builtin_fmaf(3.f, q, -r2 * q1.z), mi);
            }
        }
        nInter += (unsigned long long)__popcll(__ballot(accept));
        gd_kahan_add(tot.x, tot.cx, ax);
        gd_kahan_add(tot.y, tot.cy, ay);
        gd_kahan_add(tot.z, tot.cz, az);
        if (PHI) gd_kahan_add(phs, phc, ph);
        __builtin_amdgcn_wave_barrier();
    }
    float sx = tot.x - tot.cx, sy = tot.y - tot.cy, sz = tot.z - tot.cz, sp = phs - phc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
        sz += __shfl_xor(sz, off);
        if (PHI) sp += __shfl_xor(sp, off);
    }
    if (lane == 0) {
        acc[i] = make_float3(G * sx, G * sy, G * sz);
        if (PHI) phi[i] = -G * sp;
        if (stats) {
            unsigned long long* sl = gd_stat_slot(stats);
            atomicAdd(&sl[0], nInter);
            atomicAdd(&sl[1], nOpen);
            atomicAdd(&sl[3], nCell);
        }
    }
}

/* The per-body walk over `count` slots (or countScale * *countDev of them,
 * grid-stride, when the count is held on the device). */
template <int MODE, bool PHI, int WAVES, int STACK>
__global__ __launch_bounds__(64 * WAVES) void k_walk_body(
    const float4* __restrict__ pos, const float4* __restrict__ wcom,
    const int2* __restrict__ wlink, const float* __restrict__ wrad,
    const float4* __restrict__ wmom,
    const gd_tree_info* __restrict__ info,
    const int2* __restrict__ slotRange, const int* __restrict__ groupList,
    const int* __restrict__ bodyList, long long count,
    int* __restrict__ overflowBodies, int* __restrict__ overflowCount,
    float3* __restrict__ acc, float* __restrict__ phi, float eps2, float G, float theta,
    int n, unsigned long long* __restrict__ stats, const int* __restrict__ countDev = nullptr,
    int countScale = 1) {
    static_assert(MODE == 0 || MODE == 1, "per-body walk: monopole or quadrupole");
    __shared__ int sStackAll[WAVES][STACK];
    const int wave = WAVES > 1 ? (int)(threadIdx.x >> 6) : 0;
    int* sStack = sStackAll[wave];
    /* with countDev the count is countScale * *countDev (at most `count`); the
       grid is then fixed and each wavefront takes every stride-th body */
    long long todo = count;
    if (countDev) {
        const long long d = (long long)countScale * *countDev;
        todo = d < count ? d : count;
    }
    const long long stride = (long long)gridDim.x * WAVES;
    for (long long q = (long long)gd_xcd_block() * WAVES + wave; q < todo; q += stride) {
        long long i;
        if (bodyList) {
            i = bodyList[q];
        } else {
            const int2 rg = slotRange[groupList[q >> 6]];
            if ((int)(q & 63) >= rg.y) continue;   /* wave-uniform */
            i = rg.x + (q & 63);
        }
        if (i >= n) continue;   /* wave-uniform */
        gd_walk_one_body<MODE, PHI, STACK>(pos, wcom, wlink, wrad, wmom, info, i, sStack,
                                           overflowBodies, overflowCount, acc, phi, eps2, G,
                                           theta, stats);
        __builtin_amdgcn_wave_barrier();
    }
}

/*
 * Walk scratch.  counters[0..4]: interactions, opened, stack-overflow flag,
 * cell interactions, network evaluations (64-bit); status: overflow counts
 * of the first and second pass, longest lists, the first pass's requested
 * list entries.  First pass: packed layout from the previous walk's list
 * lengths (sizes), in a pool of listEntries.  Second pass (groups that did
 * not fit): entries2 entries per slot, at most GD_POOL_BYTES of slots at a
 * time, in batches.  Third: the per-body walk.
 */
struct gd_walk_status {
    int over1, over2, over3, deep1;
    int treeOverflow;      /* the build's node-array overflow flag, copied at the walk's start */
    int pad;
    unsigned long long requested;
};

struct gd_walk_scratch {
    unsigned long long* stats;
    gd_walk_status* status;
    int* overflow1;        /* groups entries */
    int* overflow2;        /* groups entries */
    int* bodyOver;         /* 64 x groups entries: bodies for the deep-stack per-body walk */
    int2* counts;          /* groups + 1 entries (the sentinel slot) */
    int* sizes;            /* groups entries: last list length, -1 unknown */
    long long* offsets;    /* groups entries: first-pass layout */
    int* caps;             /* groups entries */
    long long* blockSums;  /* 2 x groups / 1024 entries: layout scan */
    int2* slotRange;       /* groups + 1 entries: (first body, bodies) of each group slot */
    int* order;            /* groups entries: first-pass slot order (non-empty slots first) */
    int groupCap;          /* groups the arrays above hold */
    int* lists;            /* first-pass pool */
    long long listEntries;
    int2* counts2;         /* second-pass slots */
    int* slotCap;          /* second-pass slots: list limit per group */
    int countSlots;
    int* pool;             /* second-pass lists */
    long long poolEntries;
    int fixedCap;          /* > 0: first-pass capacity of every group (no packing) */
    int entries2;          /* second-pass list capacity */
    long long walks;       /* walks so far (spread marks are cleared every 32) */
    int deepStack;         /* first pass with 2 x GD_TR_STACK stack entries */
    unsigned long long* kstats;   /* the list passes' statistics, or NULL (not asked for) */
    float4* segPart;       /* GD_SEG_WAVES x 64: per-body sums of second-pass segments */
    int async;             /* 1: no host synchronisation inside a walk (gd_walk_set_async) */
    gd_walk_status* hstatus;   /* pinned copy of the last asynchronous walk's status */
    hipEvent_t statusEv;   /* recorded after that copy */
    int statusPending;     /* hstatus not read yet */
    gd_walk_status* hstatus1;  /* synchronous walks: pinned copy of the status after the
                                  first pass's traversal */
    hipEvent_t status1Ev;  /* recorded after that copy */
    hipEvent_t sideEv;     /* end of a synchronous walk's passes 2 and 3 (side stream) */
};

#define GD_LIST2 262144
#define GD_SEG_WAVES 2048         /* second pass: groups x segments per launch */
static_assert(GD_SEG_WAVES == 2048, "gd_seg_count assumes 2048 segment wavefronts");
#define GD_POOL_BYTES (8ll << 30)
#define GD_ASYNC_PASS2 256        /* asynchronous mode: second-pass groups per walk (the rest
                                     take the per-body walk) */
#define GD_ASYNC_BODY_BLOCKS 1024 /* asynchronous mode: per-body walk grid (grid-stride) */
#define GD_ASYNC_DEEP_BLOCKS 256  /* asynchronous mode: deep-stack per-body rerun grid */
static_assert(GD_ASYNC_PASS2 <= GD_SEG_WAVES / 2, "every asynchronous second-pass list is cut");
static gd_walk_scratch g_walk = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                 nullptr, 0,
                                 nullptr, nullptr, 0,       nullptr, 0,       0,
                                 GD_LIST2, 0,       0,       nullptr, nullptr, 0,
                                 nullptr, nullptr, 0};

static int g_lastRuns = 0;   /* 64-body runs of the last asynchronous walk */
/* Far-field ratio of the first-pass evaluation (0: off): in the quadrupole
 * and learned walks without potential, a cell with h / R below it is summed
 * through its group's local expansion (gd_far_half). */
static float g_farRatio = 0.f;
/* Opening rule of the production walks: 0 cell size, 1 mass radius. */
static int g_openRule = 0;

const float* gd_walk_radii(const gd_builder* b) { return g_openRule == 1 ? b->wrad : nullptr; }

/* optional HIP-event timing of the walk's phases (library stream) */
static int g_walkTiming = 0;
static hipEvent_t g_walkEv[4] = {nullptr, nullptr, nullptr, nullptr};
static int g_walkTimed = 0;

__global__ void k_walk_fill(int* a, int n, int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = v;
}

static void gd_walk_ensure(int groups) {
    if (gd_graphs_hold_buffers() &&
        (!g_walk.stats || groups > g_walk.groupCap || !g_walk.lists)) {
        GD_FAIL("gd_walk: the walk's buffers cannot be allocated while a graph holds the "
                "library's buffers (walk once before capturing)");
        return;
    }
    if (!g_walk.stats) {
        GD_HIP(hipMalloc(&g_walk.stats, GD_STAT_WORDS * sizeof(unsigned long long)));
        GD_HIP(hipMalloc(&g_walk.status, sizeof(gd_walk_status)));
        GD_HIP(hipMalloc(&g_walk.segPart, (size_t)GD_SEG_WAVES * 64 * sizeof(float4)));
    }
    if (groups > g_walk.groupCap) {
        if (g_walk.overflow1) GD_HIP(hipFree(g_walk.overflow1));
        if (g_walk.overflow2) GD_HIP(hipFree(g_walk.overflow2));
        if (g_walk.bodyOver) GD_HIP(hipFree(g_walk.bodyOver));
        if (g_walk.counts) GD_HIP(hipFree(g_walk.counts));
        if (g_walk.sizes) GD_HIP(hipFree(g_walk.sizes));
        if (g_walk.offsets) GD_HIP(hipFree(g_walk.offsets));
        if (g_walk.caps) GD_HIP(hipFree(g_walk.caps));
        if (g_walk.blockSums) GD_HIP(hipFree(g_walk.blockSums));
        if (g_walk.slotRange) GD_HIP(hipFree(g_walk.slotRange));
        if (g_walk.order) GD_HIP(hipFree(g_walk.order));
        GD_HIP(hipMalloc(&g_walk.overflow1, sizeof(int) * (size_t)groups));
        GD_HIP(hipMalloc(&g_walk.overflow2, sizeof(int) * (size_t)groups));
        GD_HIP(hipMalloc(&g_walk.bodyOver, 64 * sizeof(int) * (size_t)groups));
        GD_HIP(hipMalloc(&g_walk.counts, sizeof(int2) * ((size_t)groups + 1)));
        GD_HIP(hipMalloc(&g_walk.sizes, sizeof(int) * (size_t)groups));
        GD_HIP(hipMalloc(&g_walk.offsets, sizeof(long long) * (size_t)groups));
        GD_HIP(hipMalloc(&g_walk.caps, sizeof(int) * (size_t)groups));
        GD_HIP(hipMalloc(&g_walk.blockSums,
                         2 * sizeof(long long) * (size_t)gd_div_up(groups, 1024)));
        GD_HIP(hipMalloc(&g_walk.slotRange, sizeof(int2) * ((size_t)groups + 1)));
        GD_HIP(hipMalloc(&g_walk.order, sizeof(int) * (size_t)groups));
        k_walk_fill<<<gd_div_up(groups, 256), 256, 0, gd_stream()>>>(g_walk.sizes, groups, -1);
        GD_LAUNCHED();
        g_walk.groupCap = groups;
    }
    if (!g_walk.hstatus1) {
        GD_HIP(hipHostMalloc((void**)&g_walk.hstatus1, sizeof(gd_walk_status),
                             hipHostMallocDefault));
        GD_HIP(hipEventCreateWithFlags(&g_walk.status1Ev, hipEventDisableTiming));
    }
    if (!g_walk.lists) {
        /* first guess: the minimum room for every group; grown to what walks ask for */
        const long long e = (long long)groups * GD_LIST_MIN;
        GD_HIP(hipMalloc(&g_walk.lists, (size_t)e * sizeof(int)));
        g_walk.listEntries = e;
    }
}

/* Grow the first-pass pool to hold `want` entries with room to spare. */
static void gd_walk_grow_lists(long long want) {
    if (want <= g_walk.listEntries) return;
    if (gd_graphs_hold_buffers()) return;   /* a graph addresses the pool: keep it; groups
                                               without room take the later passes */
    const long long e = want + want / 4;
    /* an earlier asynchronous walk may still read the pool */
    if (g_walk.lists) GD_HIP(hipDeviceSynchronize());
    if (g_walk.lists) GD_HIP(hipFree(g_walk.lists));
    GD_HIP(hipMalloc(&g_walk.lists, (size_t)e * sizeof(int)));
    g_walk.listEntries = e;
}

static void gd_walk_ensure_pool(int slots, int entries) {
    const long long need = (long long)slots * entries;
    if (gd_graphs_hold_buffers() && (need > g_walk.poolEntries || slots > g_walk.countSlots)) {
        GD_FAIL("gd_walk: the second pass's buffers cannot be allocated while a graph holds "
                "the library's buffers (walk once asynchronously before capturing)");
        return;
    }
    if (need > g_walk.poolEntries) {
        if (g_walk.pool) GD_HIP(hipFree(g_walk.pool));
        GD_HIP(hipMalloc(&g_walk.pool, (size_t)need * sizeof(int)));
        g_walk.poolEntries = need;
    }
    if (slots > g_walk.countSlots) {
        if (g_walk.counts2) GD_HIP(hipFree(g_walk.counts2));
        if (g_walk.slotCap) GD_HIP(hipFree(g_walk.slotCap));
        GD_HIP(hipMalloc(&g_walk.counts2, (size_t)slots * sizeof(int2)));
        GD_HIP(hipMalloc(&g_walk.slotCap, (size_t)slots * sizeof(int)));
        g_walk.countSlots = slots;
    }
}

/* Group slots of the 64-body runs (see gd_traverse.h): one wavefront per
 * run; slot 2 * runs is an empty sentinel (see k_walk_layout). */
__global__ __launch_bounds__(256) void k_walk_split(const uint64_t* __restrict__ keys,
                                                    const float4* __restrict__ pos, int n,
                                                    int runs, float ratio,
                                                    int2* __restrict__ slotRange) {
    const long long g = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= runs) return;   /* wave-uniform */
    const int j = gd_split_run(keys, pos, n, g, ratio);
    if (gd_lane() == 0) {
        if (g == 0) slotRange[2 * runs] = make_int2(n, 0);   /* the sentinel: always empty */
        const int start = (int)(g * 64);
        const int count = n - start < 64 ? n - start : 64;
        slotRange[2 * g] = make_int2(start, j ? j : count);
        slotRange[2 * g + 1] = make_int2(start + (j ? j : count), j ? count - j : 0);
    }
}

/* Asynchronous mode: first-pass overflow groups beyond the second pass's
 * capacity `cap` go to the per-body walk's list. */
__global__ void k_walk_route(const int* __restrict__ over1List, const int* __restrict__ over1,
                             int cap, int* __restrict__ over2List, int* __restrict__ over2) {
    const int n1 = *over1;
    for (int i = cap + blockIdx.x * 256 + threadIdx.x; i < n1; i += gridDim.x * 256)
        over2List[atomicAdd(over2, 1)] = over1List[i];
}

__global__ void k_walk_reset(unsigned long long* stats, gd_walk_status* st,
                             const gd_tree_info* info) {
    for (int i = threadIdx.x; i < GD_STAT_WORDS; i += blockDim.x) stats[i] = 0ull;
    if (threadIdx.x == 0) {
        st->over1 = st->over2 = st->over3 = st->deep1 = 0;
        st->treeOverflow = info->overflow;
        st->pad = 0;
        st->requested = 0ull;
    }
}

template <int MODE, bool PHI>
static void gd_launch_eval(int slots, const gd_builder* b, const float4* pos,
                           const int* groupList, const int* lists, const long long* offsets,
                           const int* caps, const int2* counts, int cap, float3* acc,
                           float* phi, float eps2, float G, int n, const gd_bf16x4* W,
                           const float* Bv, float minRatio, hipStream_t stream) {
    k_walk_eval<MODE, PHI><<<gd_div_up(slots, GD_EV_WAVES), 64 * GD_EV_WAVES, 0, stream>>>(
        pos, b->wcom, b->wlink, gd_walk_radii(b), b->wmom, b->dinfo, g_walk.slotRange, groupList,
        slots, lists, offsets, caps, counts, cap, acc, phi, eps2, G, n, W, Bv, minRatio,
        g_walk.kstats);
}

template <int MODE, bool PHI>
static void gd_walk_passes(const gd_builder* b, const uint64_t* keys, const float4* pos,
                           float3* acc, float* phi, float eps2, float G, float theta, int n,
                           const gd_bf16x4* W, const float* Bv, float minRatio, int* overOut,
                           unsigned long long* requested) {
    constexpr bool SPLIT = MODE != 0;
    const int runs = gd_div_up(n, 64);
    const int groups = 2 * runs;   /* group slots: two per 64-body run */
    const float* wrad = gd_walk_radii(b);
    gd_walk_status* st = g_walk.status;
    k_walk_split<<<gd_div_up(runs, 4), 256, 0, gd_stream()>>>(keys, pos, n, runs,
                                              gd_walk_split_ratio(theta),
                                              g_walk.slotRange);
    GD_LAUNCHED();
    /* pass 1: every group, packed lists */
    const int reset = (g_walk.walks++ % 32) == 0;
    const int lblocks = gd_div_up(groups, 1024);
    k_walk_layout<<<lblocks, 1024, 0, gd_stream()>>>(g_walk.slotRange, g_walk.sizes, groups, g_walk.fixedCap,
                                     reset, g_walk.offsets, g_walk.blockSums);
    GD_LAUNCHED();
    k_walk_place<<<lblocks, 1024, 0, gd_stream()>>>(g_walk.slotRange, g_walk.sizes, groups, g_walk.fixedCap,
                                    g_walk.listEntries,
                                    g_walk.blockSums, g_walk.offsets, g_walk.caps,
                                    g_walk.order, &st->requested);
    GD_LAUNCHED();
    /* The first pass's LDS stack: GD_TR_STACK ranges per wavefront (8
       wavefronts per SIMD), or twice that (4 per SIMD) while more than 1/16
       of the runs needed more than 7/8 of GD_TR_STACK in the last walk --
       small opening angles on deep trees.  A group that overflows either
       goes to the second pass, which walks one group per wavefront at a
       fraction of the first pass's rate: worth it for a few groups, not for
       a large share. */
#define GD_PASS1(STK)                                                                          \
    k_walk_traverse<SPLIT, GD_TR_WAVES, STK><<<gd_div_up(groups, GD_TR_WAVES),                  \
                                               64 * GD_TR_WAVES, 0, gd_stream()>>>(                             \
        pos, b->wcom, b->wlink, wrad, b->dinfo, theta, n, g_walk.slotRange, g_walk.order, groups,    \
        g_walk.offsets,                                                                        \
        g_walk.caps, 0, nullptr, g_walk.lists, g_walk.counts, g_walk.sizes, g_walk.overflow1,  \
        &st->over1, g_walk.fixedCap > 0 ? nullptr : g_walk.overflow2, &st->over2, &st->deep1,   \
        g_walk.kstats)
    if (g_walk.deepStack) GD_PASS1(2 * GD_TR_STACK);
    else GD_PASS1(GD_TR_STACK);
#undef GD_PASS1
    GD_LAUNCHED();
    if (!g_walk.async) {
        /* the status (overflow counts, deepest stack, requested room) is final
           once the first pass has traversed: its copy to the host travels
           while the evaluation runs, so the host's wait below leaves the GPU
           no gap */
        GD_HIP(hipMemcpyAsync(g_walk.hstatus1, st, sizeof(gd_walk_status), hipMemcpyDeviceToHost,
                              gd_stream()));
        GD_HIP(hipEventRecord(g_walk.status1Ev, gd_stream()));
    }
    if (g_walkTiming) GD_HIP(hipEventRecord(g_walkEv[1], gd_stream()));
    {
        /* one wavefront per run: both halves' lists side by side */
        constexpr int EVW = MODE == 2 ? 8 : 4;   /* learned: 8 share the staged weights */
        const float far = g_farRatio;
        bool farDone = false;
        if constexpr (MODE != 0 && !PHI) {
            if (far > 0.f) {
                k_walk_far<EVW, MODE == 2><<<gd_div_up(runs, EVW), 64 * EVW, 0, gd_stream()>>>(
                    pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange, runs,
                    g_walk.lists, g_walk.offsets, g_walk.caps, g_walk.counts, acc, eps2,
                    far * far, minRatio, g_walk.kstats);
                GD_LAUNCHED();
                k_walk_eval_pair<MODE, PHI, EVW, true>
                    <<<gd_div_up(runs, EVW), 64 * EVW, 0, gd_stream()>>>(
                        pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange, runs,
                        g_walk.lists, g_walk.offsets, g_walk.caps, g_walk.counts, acc, phi, eps2,
                        G, W, Bv, minRatio, g_walk.kstats);
                farDone = true;
            }
        }
        if (!farDone) {
            k_walk_eval_pair<MODE, PHI, EVW, false>
                <<<gd_div_up(runs, EVW), 64 * EVW, 0, gd_stream()>>>(
                    pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange, runs,
                    g_walk.lists, g_walk.offsets, g_walk.caps, g_walk.counts, acc, phi, eps2, G,
                    W, Bv, minRatio, g_walk.kstats);
        }
    }
    GD_LAUNCHED();
    if (g_walkTiming) GD_HIP(hipEventRecord(g_walkEv[2], gd_stream()));
    constexpr int BMODE0 = MODE == 2 ? 1 : MODE;   /* per-body walk: see pass 3 below */
    if (g_walk.async) {
        /* the same passes with every decision on the device: the second pass
           has fixed grids for GD_ASYNC_PASS2 groups and reads how many it
           has; groups beyond them, and those it gives up on, take the
           per-body walk, a grid-stride loop over a device-held count */
        const int cap2 = g_walk.entries2;
        const int S2 = GD_ASYNC_PASS2;
        gd_walk_ensure_pool(S2, cap2);
        if (gd_has_error()) return;
        k_walk_route<<<64, 256, 0, gd_stream()>>>(g_walk.overflow1, &st->over1, S2, g_walk.overflow2,
                                  &st->over2);
        GD_LAUNCHED();
        k_walk_estimate<SPLIT><<<gd_div_up(S2, GD_TR_WAVES), 64 * GD_TR_WAVES, 0, gd_stream()>>>(
            pos, b->wcom, b->wlink, wrad, b->dinfo, theta, n, g_walk.slotRange, g_walk.overflow1, S2,
            cap2, cap2 / 4 < GD_PASS2_FLOOR ? cap2 / 4 : GD_PASS2_FLOOR, g_walk.slotCap,
            &st->over1);
        GD_LAUNCHED();
        k_walk_traverse<SPLIT, 1, GD_TR2_STACK><<<S2, 64, 0, gd_stream()>>>(
            pos, b->wcom, b->wlink, wrad, b->dinfo, theta, n, g_walk.slotRange, g_walk.overflow1, S2,
            nullptr, nullptr, cap2, g_walk.slotCap, g_walk.pool, g_walk.counts2, g_walk.sizes,
            g_walk.overflow2, &st->over2, nullptr, nullptr, nullptr, g_walk.kstats, &st->over1);
        GD_LAUNCHED();
        k_walk_eval_seg<MODE, PHI><<<GD_SEG_WAVES / GD_EV_WAVES, 64 * GD_EV_WAVES, 0, gd_stream()>>>(
            pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange,
            g_walk.overflow1, S2, 0, g_walk.pool, g_walk.counts2, cap2, g_walk.segPart, eps2, W,
            Bv, minRatio, g_walk.kstats, &st->over1);
        GD_LAUNCHED();
        k_walk_eval_sum<PHI><<<gd_div_up(S2, 4), 256, 0, gd_stream()>>>(g_walk.slotRange, g_walk.overflow1, S2,
                                                         0, g_walk.counts2, g_walk.segPart, acc,
                                                         phi, G, &st->over1);
        GD_LAUNCHED();
        const long long bodies = 64ll * groups;
        k_walk_body<BMODE0, PHI, GD_BODY_WAVES, GD_BODY_STACK1>
            <<<GD_ASYNC_BODY_BLOCKS, 64 * GD_BODY_WAVES, 0, gd_stream()>>>(
                pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange, g_walk.overflow2,
                nullptr, bodies, g_walk.bodyOver, &st->over3, acc, phi, eps2, G, theta, n,
                g_walk.stats, &st->over2, 64);
        GD_LAUNCHED();
        k_walk_body<BMODE0, PHI, 1, GD_BODY_STACK><<<GD_ASYNC_DEEP_BLOCKS, 64, 0, gd_stream()>>>(
            pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange, nullptr,
            g_walk.bodyOver, bodies, nullptr, nullptr, acc, phi, eps2, G, theta, n,
            g_walk.stats, &st->over3, 1);
        GD_LAUNCHED();
        /* the status goes back without waiting: the next walk (or
           gd_walk_poll) reads it; a captured walk keeps it on the device */
        if (!gd_capturing()) {
            GD_HIP(hipMemcpyAsync(g_walk.hstatus, st, sizeof(gd_walk_status),
                                  hipMemcpyDeviceToHost, gd_stream()));
            GD_HIP(hipEventRecord(g_walk.statusEv, gd_stream()));
            g_walk.statusPending = 1;
            g_lastRuns = runs;
        }
        overOut[0] = overOut[2] = overOut[3] = -1;   /* not known on the host */
        *requested = 0;
        return;
    }
    GD_HIP(hipEventSynchronize(g_walk.status1Ev));
    gd_walk_status h = *g_walk.hstatus1;
    if (h.treeOverflow) {
        /* the traversal and evaluation kernels exited at once (basicGravity.cu:282
           stops the run on the same condition) */
        GD_FAIL("Node array overflow: the tree does not fit its %d nodes", b->maxNodes);
        return;
    }
    const int over1 = h.over1;
    overOut[0] = over1 + h.over2;
    *requested = h.requested;
    g_walk.deepStack = 16ll * h.deep1 > runs;
    /* pass 2: the groups that did not fit, with long lists and a deep stack */
    const int cap2 = g_walk.entries2;
    const long long fit = GD_POOL_BYTES / ((long long)cap2 * sizeof(int));
    const int slots = over1 < fit ? over1 : (int)fit;
    if (over1 > 0) gd_walk_ensure_pool(slots, cap2);
    if (gd_has_error()) return;
    if (over1 <= 0 && h.over2 <= 0) return;
    /* passes 2 and 3 run on the side stream, beside the first pass's
       evaluation (their groups' bodies are disjoint from its); it starts
       after the traversal (the status copy above follows it) and the library
       stream waits for it at the end */
    hipStream_t side = gd_side_stream();
    GD_HIP(hipStreamWaitEvent(side, g_walk.status1Ev, 0));
    struct Join {
        hipStream_t side;
        ~Join() {
            if (!g_walk.sideEv) GD_HIP(hipEventCreateWithFlags(&g_walk.sideEv, hipEventDisableTiming));
            GD_HIP(hipEventRecord(g_walk.sideEv, side));
            GD_HIP(hipStreamWaitEvent(gd_stream(), g_walk.sideEv, 0));
        }
    } join{side};
    for (int b0 = 0; b0 < over1; b0 += slots) {
        const int cnt = over1 - b0 < slots ? over1 - b0 : slots;
        k_walk_estimate<SPLIT><<<gd_div_up(cnt, GD_TR_WAVES), 64 * GD_TR_WAVES, 0, side>>>(
            pos, b->wcom, b->wlink, wrad, b->dinfo, theta, n, g_walk.slotRange, g_walk.overflow1 + b0,
            cnt, cap2, cap2 / 4 < GD_PASS2_FLOOR ? cap2 / 4 : GD_PASS2_FLOOR, g_walk.slotCap);
        GD_LAUNCHED();
        k_walk_traverse<SPLIT, 1, GD_TR2_STACK><<<cnt, 64, 0, side>>>(
            pos, b->wcom, b->wlink, wrad, b->dinfo, theta, n, g_walk.slotRange, g_walk.overflow1 + b0,
            cnt, nullptr, nullptr, cap2, g_walk.slotCap, g_walk.pool, g_walk.counts2, g_walk.sizes,
            g_walk.overflow2, &st->over2, nullptr, nullptr, nullptr, g_walk.kstats);
        GD_LAUNCHED();
        /* few groups: each list cut into segments over many wavefronts */
        const int segs = cnt <= GD_SEG_WAVES / 2 ? (GD_SEG_WAVES / cnt < 64 ? GD_SEG_WAVES / cnt : 64)
                                                 : 1;
        if (segs > 1) {
            k_walk_eval_seg<MODE, PHI>
                <<<gd_div_up((long long)cnt * segs, GD_EV_WAVES), 64 * GD_EV_WAVES, 0, side>>>(
                    pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange,
                    g_walk.overflow1 + b0, cnt, segs, g_walk.pool, g_walk.counts2, cap2,
                    g_walk.segPart, eps2, W, Bv, minRatio, g_walk.kstats);
            GD_LAUNCHED();
            k_walk_eval_sum<PHI><<<gd_div_up(cnt, 4), 256, 0, side>>>(g_walk.slotRange,
                                                             g_walk.overflow1 + b0, cnt, segs,
                                                             g_walk.counts2, g_walk.segPart,
                                                             acc, phi, G);
        } else {
            gd_launch_eval<MODE, PHI>(cnt, b, pos, g_walk.overflow1 + b0, g_walk.pool, nullptr,
                                      nullptr, g_walk.counts2, cap2, acc, phi, eps2, G, n, W,
                                      Bv, minRatio, side);
        }
        GD_LAUNCHED();
    }
    if (over1 > 0) {   /* the side stream's own status (the null stream would wait for both) */
        GD_HIP(hipMemcpyAsync(g_walk.hstatus1, st, sizeof(h), hipMemcpyDeviceToHost, side));
        GD_HIP(hipStreamSynchronize(side));
        h = *g_walk.hstatus1;
    }
    const int over2 = h.over2;
    overOut[2] = over2;
    if (over2 <= 0) return;
    /* pass 3: body by body, 2048-node stacks; the rare body that needs more
       is walked again with a stack that fits any tree */
    const long long bodies = 64ll * over2;
    /* In the learned walk the per-body groups take the quadrupole alone: the
       network is trained and evaluated per 64-body group, and once per body
       it would cost 64 times as much for ~1 % of the bodies (the spread
       ones), which the point-based opening test already serves better. */
    constexpr int BMODE = MODE == 2 ? 1 : MODE;
    k_walk_body<BMODE, PHI, GD_BODY_WAVES, GD_BODY_STACK1>
        <<<(unsigned)gd_div_up(bodies, GD_BODY_WAVES), 64 * GD_BODY_WAVES, 0, side>>>(
            pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange, g_walk.overflow2,
            nullptr, bodies, g_walk.bodyOver, &st->over3, acc, phi, eps2, G, theta, n,
            g_walk.stats);
    GD_LAUNCHED();
    GD_HIP(hipMemcpyAsync(g_walk.hstatus1, st, sizeof(h), hipMemcpyDeviceToHost, side));
    GD_HIP(hipStreamSynchronize(side));
    h = *g_walk.hstatus1;
    overOut[3] = h.over3;
    if (h.over3 <= 0) return;
    k_walk_body<BMODE, PHI, 1, GD_BODY_STACK><<<(unsigned)h.over3, 64, 0, side>>>(
        pos, b->wcom, b->wlink, wrad, b->wmom, b->dinfo, g_walk.slotRange, nullptr, g_walk.bodyOver,
        h.over3, nullptr, nullptr, acc, phi, eps2, G, theta, n, g_walk.stats);
    GD_LAUNCHED();
}

/* The status of the last asynchronous walk, once its copy has arrived
 * (wait = 1: wait for it): grows the first-pass pool to what that walk asked
 * for, picks the traversal stack for the next walk and reports a tree that
 * did not fit its node array.  Returns -1 after reporting an error. */
static int gd_walk_take_status(int wait) {
    if (!g_walk.statusPending) return 0;
    if (wait) {
        GD_HIP(hipEventSynchronize(g_walk.statusEv));
    } else {
        const hipError_t q = hipEventQuery(g_walk.statusEv);
        if (q == hipErrorNotReady) return 0;
        GD_HIP(q);
    }
    g_walk.statusPending = 0;
    const gd_walk_status h = *g_walk.hstatus;
    if (h.treeOverflow) {
        GD_FAIL("Node array overflow: the tree of an earlier walk did not fit its nodes");
        return -1;
    }
    g_walk.deepStack = 16ll * h.deep1 > (long long)g_lastRuns;
    gd_walk_grow_lists((long long)h.requested);
    return gd_has_error() ? -1 : 0;
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
    if (mode != 0 && !(b->quad && b->wmom)) {
        GD_FAIL("gd_walk: this walk needs the second moments (gd_set_quadrupole before the build)");
        return;
    }
    if (mode == 2 && !(b->quad == 2 && b->wmom)) {
        GD_FAIL("gd_walk_learned: the network needs the third moments "
                "(gd_set_quadrupole(tree, 2) before the build)");
        return;
    }
    const int groups = 2 * gd_div_up(starCount, 64);   /* group slots */
    if (gd_capturing() && (!g_walk.async || stats || g_walkTiming)) {
        GD_FAIL("gd_walk: a captured walk must be asynchronous, without statistics or timing "
                "(gd_walk_set_async)");
        return;
    }
    if (!gd_capturing() && gd_walk_take_status(0) != 0) return;   /* a finished
                                                                     asynchronous walk's report */
    gd_walk_ensure(groups);
    if (g_walk.fixedCap > 0) gd_walk_grow_lists((long long)groups * g_walk.fixedCap);
    if (gd_has_error()) return;
    const gd_bf16x4* W = net ? gd_mlp_fwd_tiles(net) : nullptr;
    const float* Bv = net ? gd_mlp_bias(net) : nullptr;
    k_walk_reset<<<1, 256, 0, gd_stream()>>>(g_walk.stats, g_walk.status, b->dinfo);
    /* the list passes count interactions only when asked to (their per-group
       atomics are not free); the per-body walk always reports stack overflows */
    g_walk.kstats = stats ? g_walk.stats : nullptr;
    GD_LAUNCHED();
    if (g_walkTiming) {
        if (!g_walkEv[0])
            for (int e = 0; e < 4; ++e) GD_HIP(hipEventCreate(&g_walkEv[e]));
        GD_HIP(hipEventRecord(g_walkEv[0], gd_stream()));
    }
    int over[4] = {0, 0, 0, 0};
    unsigned long long requested = 0;
#define GD_PASSES(M, P)                                                                        \
    gd_walk_passes<M, P>(b, tree->mortonOut, pos, acc, phi, eps2, G, theta, starCount, W, Bv,  \
                         minRatio, over,                                                       \
                         &requested)
    if (mode == 0) {
        if (phi) GD_PASSES(0, true); else GD_PASSES(0, false);
    } else if (mode == 1) {
        if (phi) GD_PASSES(1, true); else GD_PASSES(1, false);
    } else {
        if (phi) GD_PASSES(2, true); else GD_PASSES(2, false);
    }
#undef GD_PASSES
    if (gd_has_error()) return;
    if (g_walkTiming) {
        GD_HIP(hipEventRecord(g_walkEv[3], gd_stream()));
        g_walkTimed = 1;
    }
    /* room for what this walk asked for, for the next walk (groups that did
       not get theirs went to the second pass) */
    gd_walk_grow_lists((long long)requested);
    if (g_walk.async && stats) {
        /* statistics synchronise anyway: take this walk's status now */
        if (gd_walk_take_status(1) != 0) return;
        over[0] = g_walk.hstatus->over1 + g_walk.hstatus->over2;
        over[2] = g_walk.hstatus->over2;
    }
    if ((!g_walk.async && over[0] > 0) || stats) {
        unsigned long long raw[GD_STAT_WORDS], h[7] = {0, 0, 0, 0, 0, 0, 0};
        GD_HIP(hipMemcpy(raw, g_walk.stats, sizeof(raw), hipMemcpyDeviceToHost));
        for (int sl = 0; sl < GD_STAT_SLOTS; ++sl)
            for (int k = 0; k < 7; ++k) h[k] += raw[sl * GD_STAT_STRIDE + k];
        if (h[2]) GD_FAIL("gd_walk: a body's traversal stack exceeded %d entries", GD_BODY_STACK);
        if (!stats) return;
        stats->interactions = h[0];
        stats->opened = h[1];
        stats->overflowGroups = (unsigned long long)over[0];
        stats->cellInteractions = h[3];
        stats->bodyGroups = (unsigned long long)over[2];
        stats->networkEvals = h[4];
        stats->farCells = h[5];
        stats->farInteractions = h[6];
    }
}

extern "C" void gd_walk(const treeState* tree, const float4* pos, float3* acc, float* phi,
                        float eps2, float G, float theta, int starCount, gd_walk_stats* stats) {
    const gd_builder* b = tree && tree->builder ? gd_builder_of(tree->builder) : nullptr;
    const int mode = (b && b->quad && b->wmom) ? 1 : 0;
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

extern "C" int gd_walk_set_async(int enable) {
    const int prev = g_walk.async;
    if (enable && !g_walk.hstatus) {
        GD_HIP(hipHostMalloc((void**)&g_walk.hstatus, sizeof(gd_walk_status),
                             hipHostMallocDefault));
        GD_HIP(hipEventCreateWithFlags(&g_walk.statusEv, hipEventDisableTiming));
        if (gd_has_error()) return prev;
    }
    if (!enable && gd_walk_take_status(1) != 0) return prev;
    g_walk.async = enable ? 1 : 0;
    return prev;
}

extern "C" int gd_walk_poll(void) { return gd_walk_take_status(1); }

extern "C" int gd_walk_set_list_capacity(int entries, int secondPass) {
    const int prev = g_walk.fixedCap;
    if (entries < 0 || entries > (1 << 24) || secondPass < 0 || secondPass > (1 << 26)) {
        GD_FAIL("gd_walk_set_list_capacity: capacities %d, %d out of range", entries, secondPass);
        return prev;
    }
    g_walk.fixedCap = entries;
    g_walk.entries2 = secondPass > 0 ? secondPass : GD_LIST2;
    /* forget the list lengths and spread marks recorded so far */
    if (g_walk.sizes) {
        k_walk_fill<<<gd_div_up(g_walk.groupCap, 256), 256, 0, gd_stream()>>>(g_walk.sizes, g_walk.groupCap, -1);
        GD_LAUNCHED();
    }
    return prev;
}

extern "C" float gd_walk_set_far_field(float ratio) {
    const float prev = g_farRatio;
    if (!(ratio >= 0.f && ratio < 1.f)) {   /* NaN included */
        GD_FAIL("gd_walk_set_far_field: ratio %g outside [0, 1)", (double)ratio);
        return prev;
    }
    g_farRatio = ratio;
    return prev;
}

extern "C" int gd_walk_set_opening_rule(int rule) {
    if (rule != 0 && rule != 1) {
        GD_FAIL("gd_walk_set_opening_rule: rule %d is neither 0 (cell size) nor 1 (mass radius)",
                rule);
        return -1;
    }
    const int prev = g_openRule;
    g_openRule = rule;
    return prev;
}

/* < 0: the opening-angle rule of gd_split_ratio; else the fraction for every walk */
static float g_splitRatio = -1.f;

float gd_walk_split_ratio(float theta) {
    return g_splitRatio >= 0.f ? g_splitRatio : gd_split_ratio(theta);
}

extern "C" float gd_walk_set_group_split(float ratio) {
    const float prev = g_splitRatio;
    if (!(ratio <= 1.f)) {   /* NaN included */
        GD_FAIL("gd_walk_set_group_split: fraction %g above 1", (double)ratio);
        return prev;
    }
    g_splitRatio = ratio < 0.f ? -1.f : ratio;
    return prev;
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
