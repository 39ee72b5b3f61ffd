/*
 * gd_traverse.h -- the production walk's groups and their traversal (one
 * wavefront per group), shared by the walk and by the corrector's sample
 * generator, which must see the very interaction lists the walk evaluates.
 *
 * Groups.  Bodies are taken in runs of 64 consecutive bodies in Morton order,
 * and each run is split in two at its largest key jump (the first adjacent
 * pair whose keys differ in the highest bit): each half has its own, smaller
 * bounding box and its own interaction list, and the first-pass evaluation
 * streams the two lists side by side in one wavefront, so a run costs the
 * longer of its halves' lists.  Run g owns slots 2g and 2g+1; slot 2g+1 is
 * empty when the run is not split (all keys equal, one body, or a volume
 * rule that keeps it whole).  On the 2^21-body merger the evaluation work is
 * 0.57 of whole runs' at theta 0.7 (0.85 at theta 0.2), for about the same
 * traversal work (x1.6 at theta 0.2).  A volume rule (split only when the
 * halves' boxes hold less than a fraction r <= 1 of the run's box volume)
 * remains available through gd_walk_set_group_split.
 */
#ifndef GD_TRAVERSE_H
#define GD_TRAVERSE_H

#include "gd_internal.h"

/* The default split rule: every run with two distinct keys (a fraction
 * above 1 means "always"; see gd_split_run).  The opening angle does not
 * change it: splitting pays at every angle once the halves share a
 * wavefront. */
#define GD_SPLIT_ALWAYS 2.f
__host__ __device__ inline float gd_split_ratio(float theta) {
    (void)theta;
    return GD_SPLIT_ALWAYS;
}

/* Bounding box of bodies [start, start + count), count in 1..64, in every lane. */
__device__ __forceinline__ void gd_range_box(const float4* __restrict__ pos, long long start,
                                             int count, float& lx, float& ly, float& lz,
                                             float& hx, float& hy, float& hz) {
    const int l = (int)gd_lane();
    const float4 p = pos[start + (l < count ? l : 0)];
    lx = hx = p.x; ly = hy = p.y; lz = hz = p.z;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lx = fminf(lx, __shfl_xor(lx, off)); hx = fmaxf(hx, __shfl_xor(hx, off));
        ly = fminf(ly, __shfl_xor(ly, off)); hy = fmaxf(hy, __shfl_xor(hy, off));
        lz = fminf(lz, __shfl_xor(lz, off)); hz = fmaxf(hz, __shfl_xor(hz, off));
    }
}

/* Min / max over the lanes with `in` set (in every lane); +/-inf when none. */
__device__ __forceinline__ void gd_masked_box(bool in, float x, float y, float z, float& lx,
                                              float& ly, float& lz, float& hx, float& hy,
                                              float& hz) {
    const float inf = __builtin_inff();
    lx = in ? x : inf; ly = in ? y : inf; lz = in ? z : inf;
    hx = in ? x : -inf; hy = in ? y : -inf; hz = in ? z : -inf;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lx = fminf(lx, __shfl_xor(lx, off)); hx = fmaxf(hx, __shfl_xor(hx, off));
        ly = fminf(ly, __shfl_xor(ly, off)); hy = fmaxf(hy, __shfl_xor(hy, off));
        lz = fminf(lz, __shfl_xor(lz, off)); hz = fmaxf(hz, __shfl_xor(hz, off));
    }
}

/* Where run g splits (one wavefront): returns the first body index of its
 * second half relative to the run (1..count-1), or 0 when it stays whole.
 * keys are the sorted Morton keys. */
__device__ __forceinline__ int gd_split_run(const uint64_t* __restrict__ keys,
                                            const float4* __restrict__ pos, int n, long long g,
                                            float ratio) {
    const int l = (int)gd_lane();
    const long long start = g * 64;
    const int count = (n - start) < 64 ? (int)(n - start) : 64;
    const bool valid = l < count;
    const long long i = start + (valid ? l : 0);
    const uint64_t k = keys[i];
    const uint64_t kn = (unsigned long long)__shfl((long long)k, l + 1 < 64 ? l + 1 : 63);
    const uint64_t x = (l + 1 < count) ? (k ^ kn) : 0ull;
    int hb = x ? 64 - __clzll((long long)x) : 0;
    int mx = hb;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
    if (mx == 0 || !(ratio > 0.f)) return 0;   /* all keys equal (or one body), or no splits */
    const int j = __ffsll((long long)__ballot(hb == mx));   /* lowest such lane + 1 */
    if (ratio > 1.f) return j;                               /* every run is split */
    const float4 p = pos[i];
    float ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
    gd_masked_box(valid && l < j, p.x, p.y, p.z, ax, ay, az, bx, by, bz);
    gd_masked_box(valid && l >= j, p.x, p.y, p.z, cx, cy, cz, dx, dy, dz);
    const float va = (bx - ax) * (by - ay) * (bz - az);
    const float vb = (dx - cx) * (dy - cy) * (dz - cz);
    const float lx = fminf(ax, cx), ly = fminf(ay, cy), lz = fminf(az, cz);
    const float hx = fmaxf(bx, dx), hy = fmaxf(by, dy), hz = fmaxf(bz, dz);
    const float v = (hx - lx) * (hy - ly) * (hz - lz);
    return va + vb < ratio * v ? j : 0;
}

/* One stack entry tested against the box [l, h]. */
struct gd_probe {
    int idx, nk, first;
    bool accept, open, leaf, mac;   /* mac: the node passes the size test */
};

/* The opening test: a leaf, or b^2 < theta^2 d^2 with d the distance from the
 * box to the centre of mass c and b the node's size ext / 2^level or, with
 * useRad (the mass-radius rule), its mass radius rad. */
__device__ __forceinline__ gd_probe gd_test_node(float4 c, int2 lk, int idx, float ext,
                                                 float th2, float lx, float ly, float lz,
                                                 float hx, float hy, float hz, bool useRad,
                                                 float rad) {
    gd_probe r;
    const float size = useRad ? rad : ldexpf(ext, -(lk.y & 31));
    const float ddx = fmaxf(fmaxf(lx - c.x, c.x - hx), 0.f);
    const float ddy = fmaxf(fmaxf(ly - c.y, c.y - hy), 0.f);
    const float ddz = fmaxf(fmaxf(lz - c.z, c.z - hz), 0.f);
    const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
    r.idx = idx;
    r.leaf = lk.x < 0;
    r.mac = size * size < th2 * d2;
    r.accept = r.leaf | r.mac;   /* no branch: both are cheap */
    r.open = !r.accept;
    r.nk = r.open ? (lk.y >> 5) : 0;
    r.first = lk.x;
    return r;
}

/* Lanes below this one whose bit is set in m (the hardware's masked bit
 * count: two instructions, where a popcount of m & lanemask_lt takes four). */
__device__ __forceinline__ int gd_rank(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

/* Exclusive prefix and total over the wavefront of v in [0, 15]
 * (four bit-plane ballots). */
__device__ __forceinline__ void gd_wave_scan16(int v, unsigned long long lt, int& excl,
                                               int& total) {
    excl = 0;
    total = 0;
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
        const unsigned long long bal = __ballot((v >> bit) & 1);
        excl += __popcll(bal & lt) << bit;
        total += __popcll(bal) << bit;
    }
}

/* Exclusive prefix and total over the wavefront of v (any int), by the
 * data-parallel-primitive form of the inclusive scan: shifts by 1, 2, 4, 8
 * lanes within each 16-lane row, then the last lane of row 0 / 2 added to
 * rows 1 / 3 and the last lane of row 1 to rows 2 and 3 (lanes a shift
 * reaches from outside the row, or rows a broadcast does not write, add 0). */
__device__ __forceinline__ void gd_wave_scan_dpp(int v, int& excl, int& total) {
    int s = v;
    s += __builtin_amdgcn_update_dpp(0, s, 0x111, 0xf, 0xf, false);   /* row_shr:1 */
    s += __builtin_amdgcn_update_dpp(0, s, 0x112, 0xf, 0xf, false);   /* row_shr:2 */
    s += __builtin_amdgcn_update_dpp(0, s, 0x114, 0xf, 0xf, false);   /* row_shr:4 */
    s += __builtin_amdgcn_update_dpp(0, s, 0x118, 0xf, 0xf, false);   /* row_shr:8 */
    s += __builtin_amdgcn_update_dpp(0, s, 0x142, 0xa, 0xf, false);   /* row_bcast:15 */
    s += __builtin_amdgcn_update_dpp(0, s, 0x143, 0xc, 0xf, false);   /* row_bcast:31 */
    excl = s - v;
    total = __builtin_amdgcn_readlane(s, 63);
}

/* Nodes each lane tests per step of a traversal (the pop buffer holds
 * 64 * GD_TR_PER + 8 ints).  Four (256 nodes a step, with the first pass's
 * stack cut to 896 ranges to keep eight wavefronts per SIMD) measured 2.51 ->
 * 3.21 ms for the first pass of the 16M-body merger at theta 0.65 and 10.95 ->
 * 10.50 ms at theta 0.225: at the larger angle few steps fill 256 slots and
 * the tests of the empty ones cost more than the steps saved. */
#ifndef GD_TR_PER
#define GD_TR_PER 2
#endif

/* Traversal by one wavefront for the box [l, h]: writes the accepted nodes to
 * list[0 .. cap) (leaves / all nodes from the front, cells from the back when
 * SPLIT) unless list is NULL (counting only), and returns false as soon as
 * more than limit (<= cap) entries are accepted or the stack (stackCap
 * entries) would overflow.  *opened counts the opened nodes; *depth receives
 * the deepest stack reached (on an overflow, the depth that did not fit).
 *
 * The stack holds ranges of sibling nodes, one per opened node: its first
 * child and child count (1-8) packed as first << 3 | (count - 1), so a node
 * index must stay below 2^29 (a larger one fails the traversal; the caller's
 * later passes walk the group).  A range takes one LDS word where its nodes
 * would take up to eight, which keeps the stack small enough for eight
 * wavefronts per SIMD.  Each step pops up to 64 ranges from the top -- as
 * many as hold at most 64 PER nodes -- and expands them into pop (64 PER + 8
 * ints of LDS); lane l then tests nodes l, 64 + l, ..., so that PER node
 * loads per lane are in flight.  The walk is bound by the latency of each
 * step's chain (stack, pop, node records, test), not by its instructions:
 * more nodes per step means fewer steps.  Children of the upper nodes go on
 * top: the walk stays depth-first and the stack shallow.  wrad: the nodes'
 * mass radii for the mass-radius opening rule (RAD), unused otherwise. */
template <bool SPLIT, bool RAD, int PER>
__device__ __forceinline__ bool gd_traverse_rule(
    const float4* __restrict__ wcom, const int2* __restrict__ wlink,
    const float* __restrict__ wrad,
    const gd_tree_info* __restrict__ info, float theta, float lx, float ly, float lz, float hx,
    float hy, float hz, int* stack, int stackCap, int* pop, int* __restrict__ list, int cap,
    int limit, int* nFrontOut, int* nBackOut, unsigned long long* opened, int* depth) {
    const unsigned lane = gd_lane();
    const float ext = info->box.extent;
    const float th2 = theta * theta;
    int nFront = 0, nBack = 0, top = 1, deepest = 1;
    unsigned long long nOpen = 0;
    if (lane == 0) stack[0] = 0;   /* the root: range (0, 1) */
    while (top > 0) {
        /* pop: the top ranges (lane l reads entry top-1-l), as many as hold
           <= 64 PER nodes -- at least one, as a range holds at most 8 */
        const int avail = top < 64 ? top : 64;
        const bool has = (int)lane < avail;
        const unsigned rg = has ? (unsigned)stack[top - 1 - (int)lane] : 0u;
        const int cnt = has ? (int)(rg & 7u) + 1 : 0;
        int ex, all;
        gd_wave_scan_dpp(cnt, ex, all);
        const bool take = has && ex + cnt <= 64 * PER;
        /* the taken ranges are a prefix of the lanes: their count from a
           ballot of the raw compare masked by the first avail lanes (scalar
           shifts), and the node count nn read from the last taken lane into
           a scalar register */
        const unsigned long long availMask = avail >= 64 ? ~0ull : (1ull << avail) - 1ull;
        const int nr = __popcll(__ballot(ex + cnt <= 64 * PER) & availMask);
        const int nn = __builtin_amdgcn_readlane(ex + cnt, nr - 1);
        if (take) {
            /* all 8 slots of each range, from the last: a range shorter than 8
               writes past its end into the next ranges' slots (or the 8 ints
               of slack), which those ranges write again at a lower k, in a
               later instruction -- no per-lane trip count, no divergence */
            const int f = (int)(rg >> 3);
#pragma unroll
            for (int k = 7; k >= 0; --k) {
                pop[ex + k] = f + k;
                __builtin_amdgcn_wave_barrier();
            }
        }
        __builtin_amdgcn_wave_barrier();
        top -= nr;
        /* slot p of lane l holds node 64 p + l of the step; lanes without a
           node read the root's records (no branch) and are masked below */
        bool hv[PER];
        int id[PER];
        float4 cm[PER];
        int2 lk[PER];
        float rd[PER];
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            hv[p] = 64 * p + (int)lane < nn;
            const int pv = pop[64 * p + (int)lane];
            id[p] = hv[p] ? pv : 0;
        }
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            cm[p] = wcom[id[p]];
            lk[p] = wlink[id[p]];
            rd[p] = RAD ? wrad[id[p]] : 0.f;
        }
        __builtin_amdgcn_wave_barrier();
        /* The step's lane masks come from ballots of the raw compares (leaf,
           size test) masked by the slots holding a node (scalar shifts of
           nn), and are combined by scalar operations; a ballot of a combined
           per-lane flag would cost two vector instructions more.  The
           per-lane flags that guard the stores are the same combinations of
           the same compares. */
        gd_probe pr[PER];
        unsigned long long fbm[PER], bbm[PER], obm[PER];
        int nf[PER], nbk[PER], no[PER];
        int sumF = 0, sumB = 0, sumO = 0;
        unsigned long long wide = 0ull;
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            pr[p] = gd_test_node(cm[p], lk[p], id[p], ext, th2, lx, ly, lz, hx, hy, hz, RAD,
                                 rd[p]);
            const int left = nn - 64 * p;
            const unsigned long long hm =
                left <= 0 ? 0ull : left >= 64 ? ~0ull : (1ull << left) - 1ull;
            const unsigned long long lf = __ballot(pr[p].leaf) & hm;
            const unsigned long long ac = (__ballot(pr[p].mac) & hm) | lf;
            obm[p] = hm & ~ac;
            fbm[p] = SPLIT ? lf : ac;
            bbm[p] = SPLIT ? ac & ~lf : 0ull;
            nf[p] = __popcll(fbm[p]);
            nbk[p] = __popcll(bbm[p]);
            no[p] = __popcll(obm[p]);
            sumF += nf[p];
            sumB += nbk[p];
            sumO += no[p];
            wide |= __ballot(pr[p].first >= (1 << 29)) & obm[p];
        }
        nOpen += (unsigned long long)sumO;
        deepest = max(deepest, top + sumO);
        if (nFront + sumF + nBack + sumB > limit || top + sumO > stackCap || wide != 0ull) {
            *depth = deepest;
            return false;
        }
        if (list) {
            int f0 = nFront, b0 = nBack;
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                const bool fl = hv[p] && (SPLIT ? pr[p].leaf : (pr[p].leaf || pr[p].mac));
                const bool kc = SPLIT && hv[p] && !pr[p].leaf && pr[p].mac;
                if (fl) list[f0 + gd_rank(fbm[p])] = pr[p].idx;
                if (kc) list[cap - 1 - (b0 + gd_rank(bbm[p]))] = pr[p].idx;
                f0 += nf[p];
                b0 += nbk[p];
            }
        }
        nFront += sumF;
        nBack += sumB;
        /* children of the upper nodes (lower slots) go on top; an opened
           node's child count is its link's count field */
        int below = top;
#pragma unroll
        for (int p = PER - 1; p >= 0; --p) {
            const bool op = hv[p] && !pr[p].leaf && !pr[p].mac;
            if (op)
                stack[below + gd_rank(obm[p])] =
                    (int)(((unsigned)pr[p].first << 3) | (unsigned)((lk[p].y >> 5) - 1));
            below += no[p];
        }
        top += sumO;
        __builtin_amdgcn_wave_barrier();
    }
    *nFrontOut = nFront;
    *nBackOut = nBack;
    *opened = nOpen;
    *depth = deepest;
    return true;
}

/* The traversal with the opening rule fixed per copy (wrad NULL: cell size),
 * so the node test neither computes both sizes nor branches round the radius
 * loads. */
template <bool SPLIT>
__device__ __forceinline__ bool gd_traverse_box(
    const float4* __restrict__ wcom, const int2* __restrict__ wlink,
    const float* __restrict__ wrad,
    const gd_tree_info* __restrict__ info, float theta, float lx, float ly, float lz, float hx,
    float hy, float hz, int* stack, int stackCap, int* pop, int* __restrict__ list, int cap,
    int limit, int* nFrontOut, int* nBackOut, unsigned long long* opened, int* depth) {
    if (wrad)
        return gd_traverse_rule<SPLIT, true, GD_TR_PER>(wcom, wlink, wrad, info, theta, lx, ly, lz, hx, hy,
                                             hz, stack, stackCap, pop, list, cap, limit,
                                             nFrontOut, nBackOut, opened, depth);
    return gd_traverse_rule<SPLIT, false, GD_TR_PER>(wcom, wlink, nullptr, info, theta, lx, ly, lz, hx, hy,
                                          hz, stack, stackCap, pop, list, cap, limit, nFrontOut,
                                          nBackOut, opened, depth);
}

#endif
