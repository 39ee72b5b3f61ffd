/*
 * gd_build.hip -- octree topology and centre of mass from sorted Morton keys.
 *
 * Output: the node array of the reference's level-wise build
 * (basicGravity.cu:229-319, tree.cu:50-235), byte for byte in every field the
 * reference defines.  The reference loops over 21 levels with 4 launches,
 * a scan and 2 blocking copies per level; here the topology takes three
 * launches in total and the centre of mass one launch per level, with no host
 * round trip.
 *
 * Restated rule (tree.cu:50-78): body i starts a node at level L iff it
 * shares its level-(L-1) cell with a neighbour and begins a new level-L cell.
 * With A = levels shared with the left neighbour and B with the right one
 * (-1 where there is no neighbour), that is  max(A+1, 1) <= L <= min(max(A,B)+1, 21):
 * every body starts a contiguous range of levels, usually one or two.
 *
 *   k_build_count   per 1024-body tile and level, the node starts and the
 *                   first body that begins a new cell (wave ballots)
 *   k_build_scan    per level, exclusive scan of the tile counts, and for each
 *                   tile the first cell start after it (suffix minimum)
 *   k_build_nodes   node index = level base + rank among that level's starts
 *                   (tile offset + wave offsets + lane rank); the parent is the
 *                   node of the level above that contains the body
 *                   (tree.cu:224-231); particleCount = size of the body's
 *                   level-L cell: up to the next body that begins a new
 *                   level-L cell, among the wavefront's bodies or, for a cell
 *                   that runs past them, from the first starts of the later
 *                   wavefronts and tiles (equal to tree.cu:181-203's sibling /
 *                   parent-end rule); the
 *                   first child is the body's node one level down, if it starts
 *                   one.  Written as a 16-byte meta record per node (first
 *                   body, count with the node's octant in bits 29-31, parent,
 *                   first child).
 *   k_build_mass    one launch per level, deepest first (tree.cu:89-179, same
 *                   operation order as the oracle): one thread per node of the
 *                   level below, leaves finished from their bodies, each
 *                   internal node's centre of mass and moments summed from its
 *                   children's in LDS
 *   k_node_records  the whole 80-byte node record, child[] included
 *                   (tree.cu:232-233: a child sits in the slot of its octant),
 *                   from the meta and the walk arrays; a build runs it unless
 *                   the caller asked for the compact walk arrays only (the
 *                   engine's walk reads nothing else)
 * Level tables live in gd_tree_info on the device.
 */
#include "gd_internal.h"

/* ---------------------------------------------------------------------------
 * Shared helpers
 * ------------------------------------------------------------------------- */

/* Largest L in [-1, 21] with a >> (63-3L) == b >> (63-3L). */
__device__ __forceinline__ int gd_shared_levels(uint64_t a, uint64_t b) {
    const uint64_t x = a ^ b;
    if (x == 0ull) return GD_MAX_LEVEL;
    const int lz = __clzll((long long)x);
    if (lz == 0) return -1;
    const int L = (lz - 1) / 3;
    return L > GD_MAX_LEVEL ? GD_MAX_LEVEL : L;
}

/* Tile of sorted keys plus one halo key on each side, in LDS. */
__device__ __forceinline__ void gd_load_key_tile(const uint64_t* __restrict__ keys, int n,
                                                 long long tileBase, uint64_t* sk) {
    for (int q = threadIdx.x; q < GD_BUILD_TILE + 2; q += GD_BUILD_THREADS) {
        const long long i = tileBase + q - 1;
        sk[q] = (i >= 0 && i < n) ? keys[i] : 0ull;
    }
}

/* Levels [lo, hi] at which the body at tile slot q (global index i) starts nodes. */
__device__ __forceinline__ void gd_body_levels(const uint64_t* sk, int q, long long i, int n,
                                               int* lo, int* hi, int* B) {
    const uint64_t k = sk[q + 1];
    const int A = i > 0 ? gd_shared_levels(sk[q], k) : -1;
    const int b = (i + 1 < n) ? gd_shared_levels(k, sk[q + 2]) : -1;
    const int m = A > b ? A : b;
    *lo = A + 1 > 1 ? A + 1 : 1;
    *hi = m + 1 < GD_MAX_LEVEL ? m + 1 : GD_MAX_LEVEL;
    *B = b;
}

/* Wave-uniform bounds of the levels at which the wavefront's bodies start
 * nodes: the smallest L in [1, 21] with some lo <= L (22 if none) and the
 * largest L in [1, 21] with some hi >= L (0 if none), by binary search over
 * ballots. */
__device__ __forceinline__ int gd_wave_min_lo(int lo) {
    int a = 1, b = GD_MAX_LEVEL + 1;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (__ballot(lo <= mid)) b = mid;
        else a = mid + 1;
    }
    return a;
}
__device__ __forceinline__ int gd_wave_max_hi(int hi) {
    int a = 0, b = GD_MAX_LEVEL;
    while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (__ballot(hi >= mid)) a = mid;
        else b = mid - 1;
    }
    return a;
}

/* One wavefront's node starts per level and the first of its bodies that
 * begins a new level-L cell (n if none), lane L holding level L's values.
 * Above hi's maximum no body starts a node, and a body that begins a new
 * cell there also begins one at that maximum (lo <= hi for every body that
 * starts nodes; one that duplicates its left neighbour's key has lo = 22),
 * so the first start there is the one at the maximum. */
__device__ __forceinline__ void gd_wave_level_starts(int lo, int hi, int minLo, int maxHi,
                                                     int waveBase, int n, int* cntV,
                                                     int* firstV) {
    const int lane = (int)gd_lane();
    int c = 0, f = n;
    for (int L = minLo; L <= maxHi; ++L) {
        const unsigned long long b = __ballot(lo <= L);
        const unsigned long long s = __ballot(L <= hi) & b;
        const int cL = (int)__popcll(s);
        const int fL = b ? waveBase + (int)__builtin_ctzll(b) : n;
        if (lane == L) {
            c = cL;
            f = fL;
        }
    }
    if (minLo <= maxHi) {
        const int fm = __builtin_amdgcn_readlane(f, maxHi);
        if (lane > maxHi) f = fm;
    }
    *cntV = c;
    *firstV = f;
}

/* ---------------------------------------------------------------------------
 * Topology
 * ------------------------------------------------------------------------- */
__global__ __launch_bounds__(GD_BUILD_THREADS) void k_build_count(const uint64_t* __restrict__ keys,
                                                                  int n, uint32_t* tileCounts,
                                                                  int* tileFirst, int maxTiles) {
    constexpr int WAVES = GD_BUILD_THREADS / GD_WAVE;
    __shared__ uint64_t sk[GD_BUILD_TILE + 2];
    __shared__ uint32_t sCount[WAVES][GD_LEVELS];
    __shared__ int sFirst[WAVES][GD_LEVELS];
    const long long tileBase = (long long)blockIdx.x * GD_BUILD_TILE;
    const int wave = threadIdx.x / GD_WAVE;
    gd_load_key_tile(keys, n, tileBase, sk);
    __syncthreads();
    const int lane = (int)gd_lane();
    int cnt = 0, first = n;
    for (int k = 0; k < GD_BUILD_ITEMS; ++k) {
        const int q = k * GD_BUILD_THREADS + threadIdx.x;
        const long long i = tileBase + q;
        int lo = 99, hi = -1, B;
        if (i < n) gd_body_levels(sk, q, i, n, &lo, &hi, &B);
        const int minLo = gd_wave_min_lo(lo), maxHi = gd_wave_max_hi(hi);
        int c, f;
        gd_wave_level_starts(lo, hi, minLo, maxHi,
                             (int)tileBase + k * GD_BUILD_THREADS + wave * GD_WAVE, n, &c, &f);
        cnt += c;
        first = f < first ? f : first;
    }
    if (lane >= 1 && lane <= GD_MAX_LEVEL) {
        sCount[wave][lane] = (uint32_t)cnt;
        sFirst[wave][lane] = first;
    }
    __syncthreads();
    if (threadIdx.x >= 1 && threadIdx.x <= GD_MAX_LEVEL) {
        uint32_t s = 0;
        int f = n;
        for (int w = 0; w < WAVES; ++w) {
            s += sCount[w][threadIdx.x];
            f = sFirst[w][threadIdx.x] < f ? sFirst[w][threadIdx.x] : f;
        }
        tileCounts[(size_t)threadIdx.x * maxTiles + blockIdx.x] = s;
        tileFirst[(size_t)threadIdx.x * maxTiles + blockIdx.x] = f;
    }
}

/* Block L-1 scans level L's tile counts; levelTotals[L] gets the sum.  It also
 * gives each tile the first body after it that begins a new level-L cell
 * (tileNext, n if none): the end of a level-L cell that runs past the tile. */
__global__ __launch_bounds__(1024) void k_build_scan(const uint32_t* __restrict__ tileCounts,
                                                     uint32_t* tileOffsets,
                                                     const int* __restrict__ tileFirst,
                                                     int* tileNext, int tiles, int n,
                                                     int maxTiles, uint32_t* levelTotals) {
    const int L = blockIdx.x + 1;
    const uint32_t* in = tileCounts + (size_t)L * maxTiles;
    uint32_t* out = tileOffsets + (size_t)L * maxTiles;
    __shared__ uint32_t sWave[16];
    __shared__ int sWaveMin[16];
    __shared__ uint32_t sCarry;
    __shared__ int sMinCarry;
    const unsigned lane = gd_lane();
    const int w = threadIdx.x / GD_WAVE;
    if (threadIdx.x == 0) {
        sCarry = 0;
        sMinCarry = n;
    }
    __syncthreads();
    for (int base = 0; base < tiles; base += 1024) {
        const int i = base + threadIdx.x;
        const uint32_t v = i < tiles ? in[i] : 0u;
        /* wave inclusive scan */
        uint32_t inc = v;
#pragma unroll
        for (int off = 1; off < GD_WAVE; off <<= 1) {
            const uint32_t o = __shfl_up(inc, off);
            if (lane >= (unsigned)off) inc += o;
        }
        if (lane == GD_WAVE - 1) sWave[w] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = sCarry;
            for (int x = 0; x < 16; ++x) {
                const uint32_t c = sWave[x];
                sWave[x] = run;
                run += c;
            }
            sCarry = run;
        }
        __syncthreads();
        if (i < tiles) out[i] = sWave[w] + inc - v;
        __syncthreads();
    }
    if (threadIdx.x == 0) levelTotals[L] = sCarry;

    /* exclusive suffix minimum of tileFirst, last chunk first */
    const int* fin = tileFirst + (size_t)L * maxTiles;
    int* fout = tileNext + (size_t)L * maxTiles;
    for (int base = ((tiles - 1) / 1024) * 1024; base >= 0; base -= 1024) {
        const int i = base + threadIdx.x;
        const int v = i < tiles ? fin[i] : n;
        int inc = v;   /* minimum over this lane and the wave's later lanes */
#pragma unroll
        for (int off = 1; off < GD_WAVE; off <<= 1) {
            const int o = __shfl_down(inc, off);
            if (lane + off < (unsigned)GD_WAVE) inc = o < inc ? o : inc;
        }
        if (lane == 0) sWaveMin[w] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = sMinCarry;
            for (int x = 15; x >= 0; --x) {
                const int c = sWaveMin[x];
                sWaveMin[x] = run;
                run = c < run ? c : run;
            }
            sMinCarry = run;
        }
        __syncthreads();
        int after = __shfl_down(inc, 1);
        if (lane == GD_WAVE - 1) after = n;
        if (i < tiles) fout[i] = after < sWaveMin[w] ? after : sWaveMin[w];
        __syncthreads();
    }
}

__global__ __launch_bounds__(GD_BUILD_THREADS) void k_build_nodes(
    const uint64_t* __restrict__ keys, int n, const uint32_t* __restrict__ tileOffsets,
    const int* __restrict__ tileNext, const uint32_t* __restrict__ levelTotals, int maxTiles,
    int4* __restrict__ meta, int maxNodes, gd_tree_info* info) {
    constexpr int WAVES = GD_BUILD_THREADS / GD_WAVE;
    __shared__ uint64_t sk[GD_BUILD_TILE + 2];
    __shared__ uint32_t sRank[GD_BUILD_ITEMS][WAVES][GD_LEVELS];
    __shared__ int sEnd[GD_BUILD_ITEMS][WAVES][GD_LEVELS];
    __shared__ int sLevelOff[GD_LEVELS];
    __shared__ int sOverflow;
    const long long tileBase = (long long)blockIdx.x * GD_BUILD_TILE;
    const int wave = threadIdx.x / GD_WAVE;
    const unsigned long long lt = gd_lanemask_lt();

    if (threadIdx.x == 0) {
        int off = 1, deepest = 0;
        int counts[GD_LEVELS];
        sLevelOff[0] = 0;
        counts[0] = 1;
        for (int L = 1; L <= GD_MAX_LEVEL; ++L) {
            const int c = (int)levelTotals[L];
            sLevelOff[L] = off;
            counts[L] = c;
            off += c;
            if (c > 0) deepest = L;
        }
        const int overflow = off > maxNodes;
        sOverflow = overflow;
        if (blockIdx.x == 0) {
            for (int L = 0; L <= GD_MAX_LEVEL; ++L) {
                info->levelOffsets[L] = sLevelOff[L];
                info->levelCount[L] = counts[L];
            }
            info->deepestLevel = deepest;
            info->totalNodes = off;
            info->overflow = overflow;
            info->zeroMassNodes = 0;
            /* root, basicGravity.cu:158-169: all bodies, no parent; with two
               or more bodies its first child is the first node of level 1 */
            if (!overflow) meta[0] = make_int4(0, n, -1, n >= 2 ? 1 : -1);
        }
    }
    gd_load_key_tile(keys, n, tileBase, sk);
    __syncthreads();
    if (sOverflow) return;

    /* pass 1: per (item, wave, level) the node starts and the first body that
       begins a new level-L cell (n if none) */
    int lo[GD_BUILD_ITEMS], hi[GD_BUILD_ITEMS], Bv[GD_BUILD_ITEMS];
    int minLo[GD_BUILD_ITEMS], maxHi[GD_BUILD_ITEMS];
    const int lane = (int)gd_lane();
    for (int k = 0; k < GD_BUILD_ITEMS; ++k) {
        const int q = k * GD_BUILD_THREADS + threadIdx.x;
        const long long i = tileBase + q;
        lo[k] = 99; hi[k] = -1; Bv[k] = -1;
        if (i < n) gd_body_levels(sk, q, i, n, &lo[k], &hi[k], &Bv[k]);
        minLo[k] = gd_wave_min_lo(lo[k]);
        maxHi[k] = gd_wave_max_hi(hi[k]);
        int c, f;
        gd_wave_level_starts(lo[k], hi[k], minLo[k], maxHi[k],
                             (int)tileBase + k * GD_BUILD_THREADS + wave * GD_WAVE, n, &c, &f);
        if (lane >= 1 && lane <= GD_MAX_LEVEL) {
            sRank[k][wave][lane] = (uint32_t)c;
            sEnd[k][wave][lane] = f;
        }
    }
    __syncthreads();
    if (threadIdx.x >= 1 && threadIdx.x <= GD_MAX_LEVEL) {
        /* exclusive prefix over (item, wave) in body order, from the tile's offset */
        const int L = threadIdx.x;
        uint32_t run = tileOffsets[(size_t)L * maxTiles + blockIdx.x];
        for (int k = 0; k < GD_BUILD_ITEMS; ++k)
            for (int w = 0; w < WAVES; ++w) {
                const uint32_t c = sRank[k][w][L];
                sRank[k][w][L] = run;
                run += c;
            }
    } else if (threadIdx.x >= GD_WAVE + 1 && threadIdx.x <= GD_WAVE + GD_MAX_LEVEL) {
        /* exclusive suffix minimum in body order, from the first cell start
           after the tile: sEnd becomes the end of a level-L cell that runs
           past the (item, wave)'s last body */
        const int L = threadIdx.x - GD_WAVE;
        int e = tileNext[(size_t)L * maxTiles + blockIdx.x];
        for (int k = GD_BUILD_ITEMS - 1; k >= 0; --k)
            for (int w = WAVES - 1; w >= 0; --w) {
                const int f = sEnd[k][w][L];
                sEnd[k][w][L] = e;
                e = f < e ? f : e;
            }
    }
    __syncthreads();

    /* pass 2: write the nodes, level by level over the item's wave-uniform
       range of levels with node starts */
    for (int k = 0; k < GD_BUILD_ITEMS; ++k) {
        const int L0 = minLo[k], L1 = maxHi[k];
        if (L0 > L1) continue;
        const int q = k * GD_BUILD_THREADS + threadIdx.x;
        const long long i = tileBase + q;
        const uint64_t key = sk[q + 1];
        /* no body of the wavefront starts a node at L0 - 1: each one's
           level-(L0-1) node is the last one started before the wavefront */
        int prevRank = L0 >= 2 ? (int)sRank[k][wave][L0 - 1] : 0, prevFlag = 0;
        bool flag = lo[k] <= L0 && L0 <= hi[k];
        unsigned long long fb = __ballot(flag);
        int rank = (int)sRank[k][wave][L0] + __popcll(fb & lt);
        for (int L = L0; L <= L1; ++L) {
            /* the body's node one level down, if any, is this node's first child */
            const bool flagNext = lo[k] <= L + 1 && L + 1 <= hi[k];
            unsigned long long fbNext = 0ull;
            int rankNext = 0;
            if (L + 1 <= L1) {   /* wave-uniform */
                fbNext = __ballot(flagNext);
                rankNext = (int)sRank[k][wave][L + 1] + __popcll(fbNext & lt);
            }
            if (fb) {   /* wave-uniform: some body of the wavefront starts a level-L node */
                /* level-L cells end where a body begins a new level-L cell */
                const unsigned long long ends = __ballot(lo[k] <= L);
                const unsigned long long later = ends & ~(lt | (1ull << lane));
                if (flag) {
                    const int idx = sLevelOff[L] + rank;
                    const int parent = (L == 1) ? 0 : sLevelOff[L - 1] + prevRank + prevFlag - 1;
                    const uint64_t cell = key >> (63 - 3 * L);
                    /* the cell's size: 1 when the right neighbour is outside it,
                       from the next cell start among this wavefront's bodies, or
                       from the first one after them */
                    const int count = (L > Bv[k]) ? 1
                                      : later ? (int)__builtin_ctzll(later) - lane
                                              : sEnd[k][wave][L] - (int)i;
                    /* count < 2^28 (8 n nodes fit an int): its top bits carry
                       the node's octant in its parent */
                    meta[idx] = make_int4((int)i, count | ((int)(cell & 7u) << 29), parent,
                                          flagNext ? sLevelOff[L + 1] + rankNext : -1);
                }
            }
            prevRank = rank;
            prevFlag = flag;
            rank = rankNext;
            flag = flagNext;
            fb = fbNext;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Centre of mass (tree.cu:89-179) -- shared by the engine pass and the
 * reference-API launcher.
 * ------------------------------------------------------------------------- */
struct gd_mass_result {
    float4 m;
    int leaf;
    int firstChild;
    int nChild;
    int zero;
};

/* childCom(j): child j's centre of mass (the engine reads the compact walk
 * array, the reference-API kernel the node records; the values are equal). */
template <typename ChildCom>
__device__ __forceinline__ gd_mass_result gd_node_mass(ChildCom childCom,
                                                       const float4* __restrict__ pos,
                                                       int first, int count, const int* child) {
    gd_mass_result r;
    r.m = make_float4(0.f, 0.f, 0.f, 0.f);
    r.zero = 0;
    r.nChild = 0;
    r.firstChild = -1;
#pragma unroll
    for (int c = 7; c >= 0; --c)
        if (child[c] != -1) {
            r.nChild++;
            r.firstChild = child[c];
        }
    r.leaf = r.nChild == 0;
    if (r.leaf) {
        float sx = 0.f, sy = 0.f, sz = 0.f, m = 0.f;
        for (int b = 0; b < count; ++b) {
            const float4 p = pos[first + b];
            sx = __builtin_fmaf(p.x, p.w, sx);
            sy = __builtin_fmaf(p.y, p.w, sy);
            sz = __builtin_fmaf(p.z, p.w, sz);
            m += p.w;
        }
        if (m > 0.f) r.m = make_float4(sx / m, sy / m, sz / m, m);
        else r.zero = 1;
    } else {
        /* children in slot order 0..7: masses first, then weighted positions */
        float4 q[8];
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            q[c] = child[c] != -1 ? childCom(child[c]) : make_float4(0.f, 0.f, 0.f, 0.f);
            if (child[c] != -1) a.w += q[c].w;
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (child[c] != -1) {
                a.x = __builtin_fmaf(q[c].x, q[c].w, a.x);
                a.y = __builtin_fmaf(q[c].y, q[c].w, a.y);
                a.z = __builtin_fmaf(q[c].z, q[c].w, a.z);
            }
        }
        if (a.w > 0.f) r.m = make_float4(a.x / a.w, a.y / a.w, a.z / a.w, a.w);
        else r.zero = 1;
    }
    return r;
}

/* Second moments of a leaf about its centre of mass c, S = sum m (x - c)(x - c)^T,
 * as two float4: (Sxx, Syy, Szz, Sxy), (Sxz, Syz, trace, 0). */
__device__ __forceinline__ void gd_leaf_moments(const float4* __restrict__ pos, int first,
                                                int count, float4 c, float4* q0, float4* q1) {
    float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
    for (int b = 0; b < count; ++b) {
        const float4 p = pos[first + b];
        const float dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
        sxx = __builtin_fmaf(p.w * dx, dx, sxx);
        syy = __builtin_fmaf(p.w * dy, dy, syy);
        szz = __builtin_fmaf(p.w * dz, dz, szz);
        sxy = __builtin_fmaf(p.w * dx, dy, sxy);
        sxz = __builtin_fmaf(p.w * dx, dz, sxz);
        syz = __builtin_fmaf(p.w * dy, dz, syz);
    }
    *q0 = make_float4(sxx, syy, szz, sxy);
    *q1 = make_float4(sxz, syz, sxx + syy + szz, 0.f);
}

/* Third moments of a leaf, O_ijk = sum m d_i d_j d_k about c, as three float4
 * (xxx yyy zzz xxy, xxz xyy yyz xzz, yzz xyz - -). */
__device__ __forceinline__ void gd_leaf_third(const float4* __restrict__ pos, int first,
                                              int count, float4 c, float4* o0, float4* o1,
                                              float4* o2) {
    float xxx = 0.f, yyy = 0.f, zzz = 0.f, xxy = 0.f, xxz = 0.f;
    float xyy = 0.f, yyz = 0.f, xzz = 0.f, yzz = 0.f, xyz = 0.f;
    for (int b = 0; b < count; ++b) {
        const float4 p = pos[first + b];
        const float dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
        const float mx = p.w * dx, my = p.w * dy, mz = p.w * dz;
        xxx = __builtin_fmaf(mx * dx, dx, xxx);
        yyy = __builtin_fmaf(my * dy, dy, yyy);
        zzz = __builtin_fmaf(mz * dz, dz, zzz);
        xxy = __builtin_fmaf(mx * dx, dy, xxy);
        xxz = __builtin_fmaf(mx * dx, dz, xxz);
        xyy = __builtin_fmaf(mx * dy, dy, xyy);
        yyz = __builtin_fmaf(my * dy, dz, yyz);
        xzz = __builtin_fmaf(mx * dz, dz, xzz);
        yzz = __builtin_fmaf(my * dz, dz, yzz);
        xyz = __builtin_fmaf(mx * dy, dz, xyz);
    }
    *o0 = make_float4(xxx, yyy, zzz, xxy);
    *o1 = make_float4(xxz, xyy, yyz, xzz);
    *o2 = make_float4(yzz, xyz, 0.f, 0.f);
}

/*
 * Centre of mass and moments: one launch per level L, deepest first.  Launch
 * L finishes every node of level L and writes each one's walk arrays once,
 * in index order, so consecutive threads fill whole 64-byte memory sectors
 * (a sector written in parts by different launches costs a read-modify-write
 * in memory, about five times the time of a whole one).
 *
 * A block takes a tile of 256 consecutive nodes of level L + 1 (the children
 * of level L's internal nodes), one per thread, and finishes the parents
 * whose first child lies in the tile's first 248 (a parent has at most 8
 * children, so all of theirs are in the tile; the next tile starts 248
 * further on).  Children of one parent are consecutive and in octant order.
 *   1. each thread loads its child's centre of mass and moments (the
 *      previous launch wrote them); the loads of a tile are coalesced.
 *   2. the first thread of each run of equal parents sums its children's
 *      masses, then the weighted positions, in octant order (tree.cu:127-179)
 *      -> the parent's centre of mass c, handed to each child of the run.
 *   3. each child thread forms its terms of the parent's moments about c
 *      (parallel-axis shifts: S += S_c + m d d^T, O += O_c + (S_c,ij d_k +
 *      S_c,ik d_j + S_c,jk d_i) + m d_i d_j d_k, d = c_child - c), and the
 *      run's first thread adds them in octant order.
 *   4. the tile owns the level-L nodes from its first parent up to the next
 *      tile's first parent (the first tile from the level's start, the last
 *      to its end): their internal nodes take the results of step 3 (found
 *      through their first child), their leaves are computed from their
 *      bodies (tree.cu:105-125); all are written in index order.
 * A level without children (the deepest, or the root of one body) is all
 * leaves, spread over the blocks in chunks of 256.
 * Every sum is formed term by term in the oracle's order; an octant without
 * a child adds exactly +0 in the oracle and no partial sum is ever -0 (each
 * starts at +0 and adds products or fmaf results whose addend is +0 or a sum
 * of the same kind), so leaving the empty octants out changes no bit.
 * A single-body leaf has zero moments; its moment record is not written.
 *
 * MOM: 0 centre of mass only, 1 also second moments, 2 also third moments.
 * Every build also writes each node's mass radius wrad, a bound on the
 * distance from its centre of mass to its bodies: the largest such distance
 * for a leaf (0 for one body), max over the children of |c_child - c| +
 * wrad_child for an internal node (the opening rule of
 * gd_walk_set_opening_rule).
 * Input: the per-node meta of k_build_nodes (first body, count, parent,
 * first child or -1).  Output: the walk arrays of level L.
 */
#define GD_MASS_TILE 256
#define GD_MASS_STEP (GD_MASS_TILE - 8)    /* first children per tile */

/* Finishes the level-L nodes [x0, x1) (step 4): leaves from their bodies,
 * internal nodes from the tile's results at the tile slot h of their first
 * child (tileBase = that level's index of slot 0): centre of mass sC[h],
 * children sNc[h], moment sums sTerm[e * 256 + h], mass radius sRad[h]. */
template <int MOM>
__device__ __forceinline__ int gd_finish_range(int x0, int x1, int level,
                                               const float4* __restrict__ pos,
                                               const int4* __restrict__ meta, long long tileBase,
                                               const float4* sC, const float* sTerm,
                                               const float* sRad, const int* sNc,
                                               float4* __restrict__ wcom,
                                               int2* __restrict__ wlink,
                                               float4* __restrict__ wmom,
                                               float* __restrict__ wrad) {
    int zeros = 0;
    /* a long range (leaf-heavy levels) is walked with the next chunk's meta
       in flight while this chunk's leaves load their bodies */
    int4 mtNext = make_int4(0, 0, 0, -1);
    if (x0 + (int)threadIdx.x < x1) mtNext = meta[x0 + (int)threadIdx.x];
    for (int xb = x0; xb < x1; xb += GD_MASS_TILE) {
        const int X = xb + (int)threadIdx.x;
        const int4 mt = mtNext;
        if (X + GD_MASS_TILE < x1) mtNext = meta[X + GD_MASS_TILE];
        if (X >= x1) continue;
        if (mt.w >= 0) {   /* internal: its first child's tile slot */
            const int h = (int)(mt.w - tileBase);
            const float4 c = sC[h];
            wcom[X] = c;
            wlink[X] = make_int2(mt.w, (sNc[h] << 5) | level);
            wrad[X] = sRad[h];
            if (MOM >= 1) {
                float v[16];
#pragma unroll
                for (int e = 0; e < 16; ++e)
                    v[e] = (MOM >= 2 || e < 6) ? sTerm[e * GD_MASS_TILE + h] : 0.f;
                wmom[4 * X] = make_float4(v[0], v[1], v[2], v[3]);
                wmom[4 * X + 1] = make_float4(v[4], v[5], v[6], v[7]);
                wmom[4 * X + 2] = make_float4(v[8], v[9], v[10], v[11]);
                wmom[4 * X + 3] = make_float4(v[12], v[13], v[14], v[15]);
            }
            zeros += c.w > 0.f ? 0 : 1;
        } else {           /* leaf */
            const int first = mt.x, count = mt.y & 0x1fffffff;
            float sx = 0.f, sy = 0.f, sz = 0.f, m = 0.f;
            for (int b = 0; b < count; ++b) {
                const float4 x = pos[first + b];
                sx = __builtin_fmaf(x.x, x.w, sx);
                sy = __builtin_fmaf(x.y, x.w, sy);
                sz = __builtin_fmaf(x.z, x.w, sz);
                m += x.w;
            }
            float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
            if (m > 0.f) c = make_float4(sx / m, sy / m, sz / m, m);
            else zeros++;
            wcom[X] = c;
            wlink[X] = make_int2(-(first + 1), (count << 5) | level);
            float rad = 0.f;   /* one body: its own centre of mass */
            for (int b = 0; count > 1 && b < count; ++b) {
                const float4 x = pos[first + b];
                const float dx = x.x - c.x, dy = x.y - c.y, dz = x.z - c.z;
                rad = fmaxf(rad, sqrtf((dx * dx + dy * dy) + dz * dz));
            }
            wrad[X] = rad;
            if (MOM >= 1 && count > 1) {   /* a single-body leaf's moments are zero */
                float4 q0, q1, o0 = make_float4(0.f, 0.f, 0.f, 0.f), o1 = o0, o2 = o0;
                gd_leaf_moments(pos, first, count, c, &q0, &q1);
                if (MOM >= 2) gd_leaf_third(pos, first, count, c, &o0, &o1, &o2);
                gd_mom_store(wmom, X, q0, q1.x, q1.y, o0, o1, o2);
            }
        }
    }
    return zeros;
}

template <int MOM>
__global__ __launch_bounds__(GD_MASS_TILE) __attribute__((amdgpu_waves_per_eu(6)))
void k_build_mass(const float4* __restrict__ pos,
                                                             const int4* __restrict__ meta,
                                                             gd_tree_info* info, int level,
                                                             float4* __restrict__ wcom,
                                                             int2* __restrict__ wlink,
                                                             float4* __restrict__ wmom,
                                                             float* __restrict__ wrad) {
    constexpr int TERMS = MOM == 2 ? 16 : (MOM == 1 ? 6 : 1);
    __shared__ float4 sCom[GD_MASS_TILE];     /* children's centres of mass */
    __shared__ float4 sC[GD_MASS_TILE];       /* their parent's */
    __shared__ float sTerm[TERMS * GD_MASS_TILE];   /* [term][child] */
    /* once the parents' centres of mass are formed, sCom holds the children's
       radius terms, then each run's radius at its first child */
    float* sRad = reinterpret_cast<float*>(sCom);
    __shared__ int sPar[GD_MASS_TILE + 1];    /* parent of tile child t - 1 */
    __shared__ int sNc[GD_MASS_TILE];         /* a run's children, at its first child */
    __shared__ int sRange[2];
    if (info->overflow) return;
    const int cl = level + 1;
    const int cnt = cl <= GD_MAX_LEVEL ? info->levelCount[cl] : 0;
    const int coff = cl <= GD_MAX_LEVEL ? info->levelOffsets[cl] : 0;
    const int lvStart = info->levelOffsets[level];
    const int lvEnd = lvStart + info->levelCount[level];
    const int tid = (int)threadIdx.x;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    int zeros = 0;
    if (cnt == 0) {   /* no children: every node of the level is a leaf */
        for (long long x0 = lvStart + (long long)blockIdx.x * GD_MASS_TILE; x0 < lvEnd;
             x0 += (long long)gridDim.x * GD_MASS_TILE) {
            const int x1 = x0 + GD_MASS_TILE < lvEnd ? (int)x0 + GD_MASS_TILE : lvEnd;
            zeros += gd_finish_range<MOM>((int)x0, x1, level, pos, meta, 0, sC, sTerm, sRad,
                                          sNc, wcom, wlink, wmom, wrad);
        }
        if (zeros) atomicAdd(&info->zeroMassNodes, zeros);
        return;
    }
    const long long stride = (long long)gridDim.x * GD_MASS_STEP;
    for (long long t0 = (long long)blockIdx.x * GD_MASS_STEP; t0 < cnt; t0 += stride) {
        const long long i = t0 + tid;
        const bool valid = i < cnt;
        const int j = coff + (int)i;
        /* 1. the child */
        int par = -2;
        bool hasMom = false;
        float4 ck = z4, m0 = z4, m1 = z4, m2 = z4, m3 = z4;
        float rk = 0.f;
        if (valid) {
            const int4 mt = meta[j];
            ck = wcom[j];
            rk = wrad[j];
            par = mt.z;
            hasMom = mt.w >= 0 || (mt.y & 0x1fffffff) > 1;
            if (MOM >= 1 && hasMom) {
                m0 = wmom[4 * j];
                m1 = wmom[4 * j + 1];
                if (MOM >= 2) {
                    m2 = wmom[4 * j + 2];
                    m3 = wmom[4 * j + 3];
                }
            }
        }
        sPar[tid + 1] = par;
        if (tid == 0) sPar[0] = (valid && i > 0) ? meta[j - 1].z : -3;
        sCom[tid] = ck;
        __syncthreads();
        /* 2. the first child of each run: the parent's centre of mass */
        const bool head = valid && tid < GD_MASS_STEP && sPar[tid] != par;
        int nc = 0;
        float4 c = z4;
        if (head) {
            nc = 1;
            while (nc < 8 && sPar[tid + nc + 1] == par) ++nc;
            float4 a = z4;
            for (int k = 0; k < nc; ++k) a.w += sCom[tid + k].w;
            for (int k = 0; k < nc; ++k) {
                const float4 v = sCom[tid + k];
                a.x = __builtin_fmaf(v.x, v.w, a.x);
                a.y = __builtin_fmaf(v.y, v.w, a.y);
                a.z = __builtin_fmaf(v.z, v.w, a.z);
            }
            c = a.w > 0.f ? make_float4(a.x / a.w, a.y / a.w, a.z / a.w, a.w) : z4;
            for (int k = 0; k < nc; ++k) sC[tid + k] = c;
        }
        if (tid == 0) {
            /* the owned range of level L: from the parent of the tile's first
               run (any 8 consecutive children hold a first child) to that of
               the first run at slot 248 on */
            int a = lvEnd, b = lvEnd;
            if (t0 == 0) a = lvStart;
            else
                for (int s = 0; s < 8 && t0 + s < cnt; ++s)
                    if (sPar[s] != sPar[s + 1]) { a = sPar[s + 1]; break; }
            for (int s = GD_MASS_STEP; s < GD_MASS_TILE && t0 + s < cnt; ++s)
                if (sPar[s] != sPar[s + 1]) { b = sPar[s + 1]; break; }
            sRange[0] = a;
            sRange[1] = b;
        }
        __syncthreads();
        /* 3. each child's terms about its parent's centre of mass (a child
           whose run another tile finishes writes unused ones): the radius
           bound, and with MOM the moments' parallel-axis shifts */
        {
            const float4 cp = sC[tid];
            const float dx = ck.x - cp.x, dy = ck.y - cp.y, dz = ck.z - cp.z;
            sRad[tid] = sqrtf((dx * dx + dy * dy) + dz * dz) + rk;
        }
        if (MOM >= 1) {
            {
                const float4 cp = sC[tid];
                const float4 qa = m0;
                const float4 qb = make_float4(m1.x, m1.y, 0.f, 0.f);
                const float w = ck.w;
                const float dx = ck.x - cp.x, dy = ck.y - cp.y, dz = ck.z - cp.z;
                float t[TERMS];
                t[0] = __builtin_fmaf(w * dx, dx, qa.x);
                t[1 % TERMS] = __builtin_fmaf(w * dy, dy, qa.y);
                t[2 % TERMS] = __builtin_fmaf(w * dz, dz, qa.z);
                t[3 % TERMS] = __builtin_fmaf(w * dx, dy, qa.w);
                t[4 % TERMS] = __builtin_fmaf(w * dx, dz, qb.x);
                t[5 % TERMS] = __builtin_fmaf(w * dy, dz, qb.y);
                if (MOM >= 2) {
                    /* the child's third moments: (xxx yyy) in m1.zw, (zzz xxy xxz xyy)
                       in m2, (yyz xzz yzz xyz) in m3 */
                    const float mx = w * dx, my = w * dy, mz = w * dz;
                    t[6 % TERMS] = m1.z + 3.f * qa.x * dx + mx * dx * dx;
                    t[7 % TERMS] = m1.w + 3.f * qa.y * dy + my * dy * dy;
                    t[8 % TERMS] = m2.x + 3.f * qa.z * dz + mz * dz * dz;
                    t[9 % TERMS] = m2.y + qa.x * dy + 2.f * qa.w * dx + mx * dx * dy;
                    t[10 % TERMS] = m2.z + qa.x * dz + 2.f * qb.x * dx + mx * dx * dz;
                    t[11 % TERMS] = m2.w + qa.y * dx + 2.f * qa.w * dy + mx * dy * dy;
                    t[12 % TERMS] = m3.x + qa.y * dz + 2.f * qb.y * dy + my * dy * dz;
                    t[13 % TERMS] = m3.y + qa.z * dx + 2.f * qb.x * dz + mx * dz * dz;
                    t[14 % TERMS] = m3.z + qa.z * dy + 2.f * qb.y * dz + my * dz * dz;
                    t[15 % TERMS] = m3.w + qa.w * dz + qb.x * dy + qb.y * dx + mx * dy * dz;
                }
#pragma unroll
                for (int e = 0; e < TERMS; ++e) sTerm[e * GD_MASS_TILE + tid] = t[e];
            }
            __syncthreads();
            /* the run's sums, in octant order, into its first child's slot
               (only this run's first thread reads its children's slots) */
            if (head) {
                float sum[TERMS];
#pragma unroll
                for (int e = 0; e < TERMS; ++e) sum[e] = 0.f;
                for (int k = 0; k < nc; ++k)
#pragma unroll
                    for (int e = 0; e < TERMS; ++e) sum[e] += sTerm[e * GD_MASS_TILE + tid + k];
#pragma unroll
                for (int e = 0; e < TERMS; ++e) sTerm[e * GD_MASS_TILE + tid] = sum[e];
            }
        }
        if (MOM == 0) __syncthreads();
        if (head) {
            /* the run's radius, into its first child's slot (only this run's
               first thread reads its children's slots) */
            float r = 0.f;
            for (int k = 0; k < nc; ++k) r = fmaxf(r, sRad[tid + k]);
            sRad[tid] = r;
            sNc[tid] = nc;
        }
        __syncthreads();
        /* 4. the tile's level-L nodes */
        zeros += gd_finish_range<MOM>(sRange[0], sRange[1], level, pos, meta, coff + t0, sC,
                                      sTerm, sRad, sNc, wcom, wlink, wmom, wrad);
        __syncthreads();   /* the next tile reuses the LDS */
    }
    if (zeros) atomicAdd(&info->zeroMassNodes, zeros);
}

/*
 * The 80-byte node records (types.h:52-62) of every node, from the meta and
 * the walk arrays: path = first body's key >> (63 - 3 level), parent, first
 * body, count, child[] by octant (a child's octant is 3 bits of its first
 * body's key; the first child's comes from the node's own key, the others'
 * ride in the top bits of their meta count), level, centre of mass.  Each
 * wavefront stages 64 consecutive records in LDS and stores them as 16-byte
 * pieces in address order (a store covers 1 KB of consecutive records).
 */
__global__ __launch_bounds__(256) void k_node_records(node* __restrict__ nodes,
                                                      const uint64_t* __restrict__ keys,
                                                      const int4* __restrict__ meta,
                                                      const gd_tree_info* __restrict__ info,
                                                      const float4* __restrict__ wcom,
                                                      const int2* __restrict__ wlink) {
    __shared__ int4 sRecAll[4][64 * 5];
    __shared__ int sOff[GD_LEVELS + 1];
    if (info->overflow) return;
    if (threadIdx.x < GD_LEVELS) sOff[threadIdx.x] = info->levelOffsets[threadIdx.x];
    if (threadIdx.x == 0) sOff[GD_LEVELS] = info->totalNodes;
    __syncthreads();
    const int total = sOff[GD_LEVELS];
    const unsigned lane = gd_lane();
    const int wave = (int)(threadIdx.x >> 6);
    int4* sRec = sRecAll[wave];
    const int waves = gridDim.x * 4;
    for (int c0 = (blockIdx.x * 4 + wave) * 64; c0 < total; c0 += waves * 64) {
        const int idx = c0 + (int)lane;
        if (idx < total) {
            const int4 mt = meta[idx];
            int level = 0;
#pragma unroll
            for (int L = 1; L <= GD_MAX_LEVEL; ++L) level += idx >= sOff[L] ? 1 : 0;
            const int shift = 63 - 3 * level;
            const uint64_t key = keys[mt.x];
            const float4 c = wcom[idx];
            int child[8];
#pragma unroll
            for (int o = 0; o < 8; ++o) child[o] = -1;
            if (mt.w >= 0) {
                const int fc = mt.w;
                const int nc = wlink[idx].y >> 5;
                int oct[8];
                oct[0] = (int)((key >> (shift - 3)) & 7u);
#pragma unroll
                for (int k = 1; k < 8; ++k)
                    oct[k] = k < nc ? (int)(((unsigned)meta[fc + k].y >> 29) & 7u) : 8;
#pragma unroll
                for (int k = 0; k < 8; ++k)
#pragma unroll
                    for (int o = 0; o < 8; ++o)
                        if (oct[k] == o) child[o] = fc + k;
            }
            const uint64_t path = key >> shift;
            int4* r = sRec + 5 * lane;
            r[0] = make_int4((int)(uint32_t)path, (int)(uint32_t)(path >> 32), mt.z, mt.x);
            r[1] = make_int4(mt.y & 0x1fffffff, child[0], child[1], child[2]);
            r[2] = make_int4(child[3], child[4], child[5], child[6]);
            r[3] = make_int4(child[7], level, 0, 0);
            r[4] = make_int4(__float_as_int(c.x), __float_as_int(c.y), __float_as_int(c.z),
                             __float_as_int(c.w));
        }
        __builtin_amdgcn_wave_barrier();
        int4* base = reinterpret_cast<int4*>(nodes + c0);
        const int pieces = 5 * (total - c0 < 64 ? total - c0 : 64);
#pragma unroll
        for (int u = 0; u < 5; ++u) {
            const int e = 64 * u + (int)lane;
            if (e < pieces) base[e] = sRec[e];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

hipEvent_t gd_build_mid_event = nullptr;

static int gd_node_blocks(int maxNodes, int n, int perBlock, int maxBlocks) {
    /* every node at once: at most maxNodes (and 8 n) of them */
    int blocks = gd_div_up(maxNodes < 8ll * n ? maxNodes : 8ll * n, perBlock);
    if (blocks > maxBlocks) blocks = maxBlocks;   /* the kernels stride */
    return blocks < 1 ? 1 : blocks;
}

void gd_build_node_records(gd_builder* b, const uint64_t* sortedKeys, node* nodes, int maxNodes) {
    const int n = b->lastBodies;
    if (n <= 0) return;
    k_node_records<<<gd_node_blocks(maxNodes, n, 256, 8192), 256, 0, gd_stream()>>>(
        nodes, sortedKeys, b->meta, b->dinfo, b->wcom, b->wlink);
    GD_LAUNCHED();
}

void gd_build_octree(gd_builder* b, const uint64_t* sortedKeys, const float4* sortedPos, int n,
                     node* nodes, int maxNodes, int records) {
    if (n <= 0) return;
    const int tiles = gd_div_up(n, GD_BUILD_TILE);
    if (tiles > b->maxTiles) {
        GD_FAIL("octree build of %d bodies exceeds the builder's capacity", n);
        return;
    }
    b->lastBodies = n;
    k_build_count<<<tiles, GD_BUILD_THREADS, 0, gd_stream()>>>(sortedKeys, n, b->tileCounts,
                                                               b->tileFirst, b->maxTiles);
    GD_LAUNCHED();
    k_build_scan<<<GD_MAX_LEVEL, 1024, 0, gd_stream()>>>(b->tileCounts, b->tileOffsets, b->tileFirst,
                                                         b->tileNext, tiles, n, b->maxTiles,
                                                         b->levelTotals);
    GD_LAUNCHED();
    k_build_nodes<<<tiles, GD_BUILD_THREADS, 0, gd_stream()>>>(sortedKeys, n, b->tileOffsets,
                                                               b->tileNext, b->levelTotals,
                                                               b->maxTiles, b->meta, maxNodes,
                                                               b->dinfo);
    GD_LAUNCHED();
    if (gd_build_mid_event) GD_HIP(hipEventRecord(gd_build_mid_event, gd_stream()));
    /* launch L finishes level L from its children at L + 1 (level 21 holds
       leaves only; an empty level returns at once) */
    for (int L = GD_MAX_LEVEL; L >= 0; --L) {
        /* level L + 1 holds at most min(8^(L+1), n) nodes */
        long long bound = n;
        if (L + 1 < 8) {
            const long long p8 = 1ll << (3 * (L + 1));
            bound = p8 < bound ? p8 : bound;
        }
        int lb = gd_div_up(bound, GD_MASS_STEP);
        /* at most the blocks that fit at once (6, 7, 8 per CU for third,
           second, no moments); the kernel strides */
        const int fit = b->quad == 2 ? 1536 : (b->quad == 1 ? 1792 : 2048);
        if (lb > fit) lb = fit;
        if (b->quad == 2)
            k_build_mass<2><<<lb, GD_MASS_TILE, 0, gd_stream()>>>(sortedPos, b->meta, b->dinfo, L,
                                                                  b->wcom, b->wlink, b->wmom, b->wrad);
        else if (b->quad == 1)
            k_build_mass<1><<<lb, GD_MASS_TILE, 0, gd_stream()>>>(sortedPos, b->meta, b->dinfo, L,
                                                                  b->wcom, b->wlink, b->wmom, b->wrad);
        else
            k_build_mass<0><<<lb, GD_MASS_TILE, 0, gd_stream()>>>(sortedPos, b->meta, b->dinfo, L,
                                                                  b->wcom, b->wlink, nullptr, b->wrad);
        GD_LAUNCHED();
    }
    if (records) gd_build_node_records(b, sortedKeys, nodes, maxNodes);
}

/* ---------------------------------------------------------------------------
 * Reference-API level kernels (tree.cu:50-78, 181-235, 89-179), for callers
 * that drive the level loop themselves.
 * ------------------------------------------------------------------------- */
__device__ __forceinline__ uint64_t gd_cell_at(uint64_t key, int level) {
    return key >> (63 - 3 * level);
}

__global__ __launch_bounds__(256) void k_identify(const uint64_t* __restrict__ keys, int level,
                                                  int n, uint32_t* flags) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t here = gd_cell_at(keys[i], level);
    const uint64_t up = gd_cell_at(keys[i], level - 1);
    const bool upLeft = i > 0 && up == gd_cell_at(keys[i - 1], level - 1);
    const bool upRight = i < n - 1 && up == gd_cell_at(keys[i + 1], level - 1);
    const bool sameLeft = i > 0 && here == gd_cell_at(keys[i - 1], level);
    flags[i] = (!upLeft && !upRight) ? 0u : (sameLeft ? 0u : 1u);
}

__global__ __launch_bounds__(256) void k_create_nodes(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flags,
    const uint32_t* __restrict__ prevFlags, const uint32_t* __restrict__ offs,
    const uint32_t* __restrict__ prevOffs, node* nodes, int level, int base, int prevBase, int n,
    int maxNodes) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || flags[i] == 0u) return;
    const uint64_t path = gd_cell_at(keys[i], level);
    const int idx = (int)offs[i] + base;
    if (idx >= maxNodes) return;
    node* nd = nodes + idx;
    nd->nodePathFromRoot = path;
    nd->treeLevel = level;
    nd->firstParticleIndex = i;
    for (int c = 0; c < 8; ++c) nd->child[c] = -1;
    const int parent = level == 1 ? 0 : prevBase + (int)prevOffs[i] + (int)prevFlags[i] - 1;
    nodes[parent].child[path & 7u] = idx;
    nd->parentIndex = parent;
}

__global__ __launch_bounds__(256) void k_tree_particles(node* nodes, int levelStart, int levelCount) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= levelCount) return;
    const int idx = levelStart + t;
    const int parent = nodes[idx].parentIndex;
    int end;
    if (t + 1 < levelCount && nodes[idx + 1].parentIndex == parent)
        end = nodes[idx + 1].firstParticleIndex;
    else
        end = nodes[parent].firstParticleIndex + nodes[parent].particleCount;
    nodes[idx].particleCount = end - nodes[idx].firstParticleIndex;
}

__global__ __launch_bounds__(256) void k_tree_mass_api(node* nodes, const float4* __restrict__ pos,
                                                       int levelNodes, int levelStart) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= levelNodes) return;
    const int idx = levelStart + t;
    int child[8];
    for (int c = 0; c < 8; ++c) child[c] = nodes[idx].child[c];
    const gd_mass_result m = gd_node_mass([nodes](int j) { return nodes[j].massData; }, pos,
                                          nodes[idx].firstParticleIndex,
                                          nodes[idx].particleCount, child);
    nodes[idx].massData = m.m;
}

extern "C" {

void identifyNodesAtLevel(uint64_t* mortonCodes, int nodeLevel, int starCount, uint32_t* nodeFlags) {
    if (starCount <= 0) return;
    k_identify<<<gd_div_up(starCount, 256), 256, 0, gd_stream()>>>(mortonCodes, nodeLevel, starCount, nodeFlags);
    GD_LAUNCHED();
}

void createNodes(const uint64_t* mortonCodes, const uint32_t* flags, const uint32_t* previousFlags,
                 const uint32_t* offsets, const uint32_t* offsetsPrev, node* nodes, int level,
                 int arrayLevelOffset, int prevArrayLevelOffset, int starCount, int maxNodes) {
    if (starCount <= 0) return;
    k_create_nodes<<<gd_div_up(starCount, 256), 256, 0, gd_stream()>>>(mortonCodes, flags, previousFlags, offsets,
                                                       offsetsPrev, nodes, level, arrayLevelOffset,
                                                       prevArrayLevelOffset, starCount, maxNodes);
    GD_LAUNCHED();
}

void updateTreeParticles(node* nodes, int levelStartIndex, int levelNodeCount, int starCount) {
    (void)starCount;
    if (levelNodeCount <= 0) return;
    k_tree_particles<<<gd_div_up(levelNodeCount, 256), 256, 0, gd_stream()>>>(nodes, levelStartIndex,
                                                               levelNodeCount);
    GD_LAUNCHED();
}

void updateTreeMass(node* nodes, const float4* positionMassVals, int level, int levelNodes,
                    int levelStartIndex) {
    (void)level;
    if (levelNodes <= 0) return;
    k_tree_mass_api<<<gd_div_up(levelNodes, 256), 256, 0, gd_stream()>>>(nodes, positionMassVals, levelNodes,
                                                         levelStartIndex);
    GD_LAUNCHED();
}

}  /* extern "C" */
