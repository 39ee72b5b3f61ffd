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
 *   k_build_count   per 1024-body tile, node starts per level (wave ballots)
 *   k_build_scan    per level, exclusive scan of the tile counts
 *   k_build_nodes   node index = level base + rank among that level's starts
 *                   (tile offset + wave offsets + lane rank); the parent is the
 *                   node of the level above that contains the body
 *                   (tree.cu:224-231); particleCount = size of the body's
 *                   level-L cell, found by galloping over the sorted keys
 *                   (equal to tree.cu:181-203's sibling / parent-end rule); the
 *                   first child is the body's node one level down, if it starts
 *                   one.  Written as a 16-byte meta record per node (first
 *                   body, count with the node's octant in bits 29-31, parent,
 *                   first child).
 *   k_build_mass    one launch per level, deepest first (tree.cu:89-179, same
 *                   operation order as the oracle): centre of mass, moments,
 *                   the compact walk record (centre of mass, first child /
 *                   body, count, level) and the whole 80-byte node record,
 *                   child[] included (tree.cu:232-233: a child sits in the slot
 *                   of its octant)
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

/* ---------------------------------------------------------------------------
 * Topology
 * ------------------------------------------------------------------------- */
__global__ __launch_bounds__(GD_BUILD_THREADS) void k_build_count(const uint64_t* __restrict__ keys,
                                                                  int n, uint32_t* tileCounts,
                                                                  int maxTiles) {
    constexpr int WAVES = GD_BUILD_THREADS / GD_WAVE;
    __shared__ uint64_t sk[GD_BUILD_TILE + 2];
    __shared__ uint32_t sCount[WAVES][GD_LEVELS];
    const long long tileBase = (long long)blockIdx.x * GD_BUILD_TILE;
    const int wave = threadIdx.x / GD_WAVE;
    gd_load_key_tile(keys, n, tileBase, sk);
    if (threadIdx.x < WAVES * GD_LEVELS) (&sCount[0][0])[threadIdx.x] = 0u;
    __syncthreads();
    for (int k = 0; k < GD_BUILD_ITEMS; ++k) {
        const int q = k * GD_BUILD_THREADS + threadIdx.x;
        const long long i = tileBase + q;
        int lo = 99, hi = -1, B;
        if (i < n) gd_body_levels(sk, q, i, n, &lo, &hi, &B);
        for (int L = 1; L <= GD_MAX_LEVEL; ++L) {
            const unsigned long long bal = __ballot(lo <= L && L <= hi);
            if (gd_lane() == 0) sCount[wave][L] += (uint32_t)__popcll(bal);
        }
    }
    __syncthreads();
    if (threadIdx.x >= 1 && threadIdx.x <= GD_MAX_LEVEL) {
        uint32_t s = 0;
        for (int w = 0; w < WAVES; ++w) s += sCount[w][threadIdx.x];
        tileCounts[(size_t)threadIdx.x * maxTiles + blockIdx.x] = s;
    }
}

/* Block L-1 scans level L's tile counts; levelTotals[L] gets the sum. */
__global__ __launch_bounds__(1024) void k_build_scan(const uint32_t* __restrict__ tileCounts,
                                                     uint32_t* tileOffsets, int tiles,
                                                     int maxTiles, uint32_t* levelTotals) {
    const int L = blockIdx.x + 1;
    const uint32_t* in = tileCounts + (size_t)L * maxTiles;
    uint32_t* out = tileOffsets + (size_t)L * maxTiles;
    __shared__ uint32_t sWave[16];
    __shared__ uint32_t sCarry;
    if (threadIdx.x == 0) sCarry = 0;
    __syncthreads();
    for (int base = 0; base < tiles; base += 1024) {
        const int i = base + threadIdx.x;
        const uint32_t v = i < tiles ? in[i] : 0u;
        /* wave inclusive scan */
        uint32_t inc = v;
        const unsigned lane = gd_lane();
#pragma unroll
        for (int off = 1; off < GD_WAVE; off <<= 1) {
            const uint32_t o = __shfl_up(inc, off);
            if (lane >= (unsigned)off) inc += o;
        }
        if (lane == GD_WAVE - 1) sWave[threadIdx.x / GD_WAVE] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = sCarry;
            for (int w = 0; w < 16; ++w) {
                const uint32_t c = sWave[w];
                sWave[w] = run;
                run += c;
            }
            sCarry = run;
        }
        __syncthreads();
        if (i < tiles) out[i] = sWave[threadIdx.x / GD_WAVE] + inc - v;
        __syncthreads();
    }
    if (threadIdx.x == 0) levelTotals[L] = sCarry;
}

/* Cell size: first j > i whose level-L cell differs, minus i (galloping). */
__device__ __forceinline__ int gd_cell_size(const uint64_t* __restrict__ keys, int n, long long i,
                                            uint64_t cell, int shift) {
    long long same = i + 1;     /* keys[i+1] is known to share the cell */
    long long step = 2;
    long long probe = i + step;
    while (probe < n && (keys[probe] >> shift) == cell) {
        same = probe;
        step <<= 1;
        probe = i + step;
    }
    long long diff = probe < n ? probe : n;
    while (diff - same > 1) {
        const long long mid = same + ((diff - same) >> 1);
        if ((keys[mid] >> shift) == cell) same = mid; else diff = mid;
    }
    return (int)(diff - i);
}

__global__ __launch_bounds__(GD_BUILD_THREADS) void k_build_nodes(
    const uint64_t* __restrict__ keys, int n, const uint32_t* __restrict__ tileOffsets,
    const uint32_t* __restrict__ levelTotals, int maxTiles, int4* __restrict__ meta,
    int maxNodes, gd_tree_info* info) {
    constexpr int WAVES = GD_BUILD_THREADS / GD_WAVE;
    __shared__ uint64_t sk[GD_BUILD_TILE + 2];
    __shared__ uint32_t sRank[GD_BUILD_ITEMS][WAVES][GD_LEVELS];
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

    /* pass 1: per (item, wave, level) counts of node starts */
    int lo[GD_BUILD_ITEMS], hi[GD_BUILD_ITEMS], Bv[GD_BUILD_ITEMS];
    int waveMaxHi = 0;
    for (int k = 0; k < GD_BUILD_ITEMS; ++k) {
        const int q = k * GD_BUILD_THREADS + threadIdx.x;
        const long long i = tileBase + q;
        lo[k] = 99; hi[k] = -1; Bv[k] = -1;
        if (i < n) gd_body_levels(sk, q, i, n, &lo[k], &hi[k], &Bv[k]);
        for (int L = 1; L <= GD_MAX_LEVEL; ++L) {
            const unsigned long long bal = __ballot(lo[k] <= L && L <= hi[k]);
            if (gd_lane() == 0) sRank[k][wave][L] = (uint32_t)__popcll(bal);
        }
        waveMaxHi = hi[k] > waveMaxHi ? hi[k] : waveMaxHi;
    }
    __syncthreads();
    /* exclusive prefix over (item, wave) in body order, from the tile's offset */
    if (threadIdx.x >= 1 && threadIdx.x <= GD_MAX_LEVEL) {
        const int L = threadIdx.x;
        uint32_t run = tileOffsets[(size_t)L * maxTiles + blockIdx.x];
        for (int k = 0; k < GD_BUILD_ITEMS; ++k)
            for (int w = 0; w < WAVES; ++w) {
                const uint32_t c = sRank[k][w][L];
                sRank[k][w][L] = run;
                run += c;
            }
    }
    __syncthreads();

    /* pass 2: write the nodes */
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(waveMaxHi, off);
        waveMaxHi = o > waveMaxHi ? o : waveMaxHi;
    }
    for (int k = 0; k < GD_BUILD_ITEMS; ++k) {
        const int q = k * GD_BUILD_THREADS + threadIdx.x;
        const long long i = tileBase + q;
        const uint64_t key = sk[q + 1];
        int prevRank = 0, prevFlag = 0;
        bool flag = lo[k] <= 1 && 1 <= hi[k];
        int rank = (int)sRank[k][wave][1] + __popcll(__ballot(flag) & lt);
        for (int L = 1; L <= waveMaxHi; ++L) {
            /* the body's node one level down, if any, is this node's first child */
            const bool flagNext = lo[k] <= L + 1 && L + 1 <= hi[k];
            int rankNext = 0;
            if (L + 1 <= waveMaxHi)   /* wave-uniform */
                rankNext = (int)sRank[k][wave][L + 1] + __popcll(__ballot(flagNext) & lt);
            /* level-L cells end where a body starts a node at a level <= L
               (its left neighbour differs from it at level L) */
            const unsigned long long ends = __ballot(lo[k] <= L);
            if (flag) {
                const int idx = sLevelOff[L] + rank;
                const int parent = (L == 1) ? 0 : sLevelOff[L - 1] + prevRank + prevFlag - 1;
                const int shift = 63 - 3 * L;
                const uint64_t cell = key >> shift;
                /* the cell's size: 1 when the right neighbour is outside it,
                   from the next cell start among this wavefront's bodies, or
                   by galloping over the keys when the cell runs past them */
                const unsigned long long later = ends & ~(lt | (1ull << gd_lane()));
                const int count = (L > Bv[k]) ? 1
                                  : later ? (int)__builtin_ctzll(later) - (int)gd_lane()
                                          : gd_cell_size(keys, n, i, cell, shift);
                /* count < 2^28 (8 n nodes fit an int): its top bits carry
                   the node's octant in its parent */
                meta[idx] = make_int4((int)i, count | ((int)(cell & 7u) << 29), parent,
                                      flagNext ? sLevelOff[L + 1] + rankNext : -1);
            }
            prevRank = rank;
            prevFlag = flag;
            rank = rankNext;
            flag = flagNext;
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

/* A single-body leaf's moments about its centre of mass are zero and are
 * not stored: no walk reads a leaf's moments (leaves are monopoles), and a
 * parent that finds such a child (its link: a leaf of count 1) takes them as
 * zero.  Three quarters of the nodes are such leaves, so this saves most of
 * the moment traffic of the build (80 B written and read back per leaf). */
__device__ __forceinline__ bool gd_single_leaf(const int2* __restrict__ wlink, int j) {
    const int2 lk = wlink[j];
    return lk.x < 0 && (lk.y >> 5) == 1;
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
 * An internal node of the engine build: centre of mass in the oracle's
 * operation order (masses of the children in slot order, then the weighted
 * positions), then its moments about it.  Every child's centre of mass and
 * link are loaded without a branch (an empty slot reads the first child
 * again and its contribution is selected away), so all loads are in flight
 * together instead of one child's dependent chain at a time; the moment
 * records are then loaded only for children that store them (the lanes'
 * other loads are masked off, so they cost no memory traffic: single-body
 * leaves, three quarters of the nodes, would otherwise add 80 B of reads
 * each).  Second moments shift
 * each child's by the parallel-axis term m_c d d^T (d = c_child - c); third
 * moments: O += O_c + (S_c,ij d_k + S_c,ik d_j + S_c,jk d_i) + m_c d_i d_j d_k.
 * A single-body leaf child has zero moments, which are not stored.
 */
template <int MOM>
__device__ __forceinline__ float4 gd_inner_node(const float4* __restrict__ wcom,
                                                const int2* __restrict__ wlink,
                                                const float4* __restrict__ wquad,
                                                const float4* __restrict__ woct,
                                                const int* child, int fc, float4* q0,
                                                float4* q1, float4* o0, float4* o1,
                                                float4* o2) {
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 cc[8];
    int jj[8];
    bool mom[8];   /* the child stores moments: not an empty slot, not a single-body leaf */
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        jj[k] = child[k] != -1 ? child[k] : fc;
        const float4 v = wcom[jj[k]];
        cc[k] = child[k] != -1 ? v : z4;
        if (MOM != 0) mom[k] = child[k] != -1 && !gd_single_leaf(wlink, jj[k]);
    }
    float4 a = z4;
#pragma unroll
    for (int k = 0; k < 8; ++k) a.w += cc[k].w;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a.x = __builtin_fmaf(cc[k].x, cc[k].w, a.x);
        a.y = __builtin_fmaf(cc[k].y, cc[k].w, a.y);
        a.z = __builtin_fmaf(cc[k].z, cc[k].w, a.z);
    }
    const float4 c = a.w > 0.f ? make_float4(a.x / a.w, a.y / a.w, a.z / a.w, a.w) : z4;
    if (MOM == 0) return c;
    float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
    float xxx = 0.f, yyy = 0.f, zzz = 0.f, xxy = 0.f, xxz = 0.f;
    float xyy = 0.f, yyz = 0.f, xzz = 0.f, yzz = 0.f, xyz = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int j = jj[k];
        const bool use = mom[k];
        /* only children that store moments are read (a single-body leaf's
           moments are zero and were never written) */
        float4 qa = z4, qb = z4;
        if (use) {
            qa = wquad[2 * j];
            qb = wquad[2 * j + 1];
        }
        const float w = cc[k].w;   /* 0 for an empty slot */
        const float dx = cc[k].x - c.x, dy = cc[k].y - c.y, dz = cc[k].z - c.z;
        if (MOM >= 2) {
            float4 p = z4, q = z4, r = z4;
            if (use) {
                p = woct[3 * j];
                q = woct[3 * j + 1];
                r = woct[3 * j + 2];
            }
            const float mx = w * dx, my = w * dy, mz = w * dz;
            xxx += p.x + 3.f * qa.x * dx + mx * dx * dx;
            yyy += p.y + 3.f * qa.y * dy + my * dy * dy;
            zzz += p.z + 3.f * qa.z * dz + mz * dz * dz;
            xxy += p.w + qa.x * dy + 2.f * qa.w * dx + mx * dx * dy;
            xxz += q.x + qa.x * dz + 2.f * qb.x * dx + mx * dx * dz;
            xyy += q.y + qa.y * dx + 2.f * qa.w * dy + mx * dy * dy;
            yyz += q.z + qa.y * dz + 2.f * qb.y * dy + my * dy * dz;
            xzz += q.w + qa.z * dx + 2.f * qb.x * dz + mx * dz * dz;
            yzz += r.x + qa.z * dy + 2.f * qb.y * dz + my * dz * dz;
            xyz += r.y + qa.w * dz + qb.x * dy + qb.y * dx + mx * dy * dz;
        }
        sxx += __builtin_fmaf(w * dx, dx, qa.x);
        syy += __builtin_fmaf(w * dy, dy, qa.y);
        szz += __builtin_fmaf(w * dz, dz, qa.z);
        sxy += __builtin_fmaf(w * dx, dy, qa.w);
        sxz += __builtin_fmaf(w * dx, dz, qb.x);
        syz += __builtin_fmaf(w * dy, dz, qb.y);
    }
    *q0 = make_float4(sxx, syy, szz, sxy);
    *q1 = make_float4(sxz, syz, sxx + syy + szz, 0.f);
    if (MOM >= 2) {
        *o0 = make_float4(xxx, yyy, zzz, xxy);
        *o1 = make_float4(xxz, xyy, yyz, xzz);
        *o2 = make_float4(yzz, xyz, 0.f, 0.f);
    }
    return c;
}

/* One finished node record (types.h:52-62), written whole: five 16-byte
 * stores, so consecutive nodes fill whole cache lines. */
__device__ __forceinline__ void gd_write_record(node* __restrict__ nodes, int idx, uint64_t path,
                                                int parent, int first, int count, const int* child,
                                                int level, float4 com) {
    int4* r = reinterpret_cast<int4*>(nodes + idx);
    r[0] = make_int4((int)(uint32_t)path, (int)(uint32_t)(path >> 32), parent, first);
    r[1] = make_int4(count, child[0], child[1], child[2]);
    r[2] = make_int4(child[3], child[4], child[5], child[6]);
    r[3] = make_int4(child[7], level, 0, 0);
    r[4] = make_int4(__float_as_int(com.x), __float_as_int(com.y), __float_as_int(com.z),
                     __float_as_int(com.w));
}

/* One internal node at level `level` (shift = 63 - 3 level): its children
 * are the consecutive nodes of the next level from its first child on whose
 * parent it is (at most 8, before childEnd), in octant order; a child's
 * octant (3 bits of its first body's key) rides in the top bits of its meta
 * count, so no child's key is read.  Writes the walk arrays and the
 * whole node record; returns 1 for a massless node. */
template <int MOM>
__device__ __forceinline__ int gd_finish_inner(node* __restrict__ nodes,
                                               const uint64_t* __restrict__ keys,
                                               const int4* __restrict__ meta, int idx, int level,
                                               int childEnd, float4* __restrict__ wcom,
                                               int2* __restrict__ wlink,
                                               float4* __restrict__ wquad,
                                               float4* __restrict__ woct) {
    const int shift = 63 - 3 * level;
    const int4 mt = meta[idx];
    const int fc = mt.w;
    const uint64_t key = keys[mt.x];
    int kid[8], oct[8];
    kid[0] = fc;
    oct[0] = (int)((key >> (shift - 3)) & 7u);
    bool run = true;
#pragma unroll
    for (int k = 1; k < 8; ++k) {
        const int j = fc + k;
        const int4 mk = j < childEnd ? meta[j] : make_int4(0, 0, -1, -1);
        run = run && mk.z == idx;
        kid[k] = run ? j : -1;
        oct[k] = run ? (int)(((unsigned)mk.y >> 29) & 7u) : 8;
    }
    int nc = 0;
    int child[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        nc += kid[k] >= 0 ? 1 : 0;
        child[k] = -1;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int o = 0; o < 8; ++o)
            if (oct[k] == o) child[o] = kid[k];
    float4 q0, q1, o0, o1, o2;
    const float4 c =
        gd_inner_node<MOM>(wcom, wlink, wquad, woct, kid, fc, &q0, &q1, &o0, &o1, &o2);
    gd_write_record(nodes, idx, key >> shift, mt.z, mt.x, mt.y & 0x1fffffff, child, level, c);
    wcom[idx] = c;
    wlink[idx] = make_int2(fc, (nc << 5) | level);
    if (MOM >= 1) {
        wquad[2 * idx] = q0;
        wquad[2 * idx + 1] = q1;
    }
    if (MOM >= 2) {
        woct[3 * idx] = o0;
        woct[3 * idx + 1] = o1;
        woct[3 * idx + 2] = o2;
    }
    return c.w > 0.f ? 0 : 1;
}

/*
 * Engine centre of mass, moments and node records for one level (launched
 * deepest first).  MOM: 0 centre of mass only, 1 also second moments, 2 also
 * third moments.  Input: the per-node meta of k_build_nodes (first body,
 * count, parent, first child or -1).  Output: the walk arrays and the whole
 * 80-byte node record (sRec: 64 records of LDS per wavefront, for the leaves'
 * coalesced stores).  Each wavefront scans 64 nodes at a time: a lane whose
 * node is a leaf finishes it at once (tree.cu:105-125; three quarters of all
 * nodes are single-body leaves, which store no moments), the internal nodes
 * (tree.cu:127-179) are queued in the wavefront's LDS and finished 64 at a
 * time with all lanes, so the long internal path runs in full wavefronts and
 * no block-wide barrier holds a wavefront back.
 */
template <int MOM>
__global__ __launch_bounds__(256) void k_build_mass(node* __restrict__ nodes,
                                                    const float4* __restrict__ pos,
                                                    const uint64_t* __restrict__ keys,
                                                    const int4* __restrict__ meta,
                                                    gd_tree_info* info, int level,
                                                    float4* __restrict__ wcom,
                                                    int2* __restrict__ wlink,
                                                    float4* __restrict__ wquad,
                                                    float4* __restrict__ woct) {
    __shared__ int sQueueAll[4][128];
    __shared__ int4 sRecAll[4][64 * 5];
    if (info->overflow) return;
    const int cnt = info->levelCount[level];
    const int off = info->levelOffsets[level];
    const int childEnd =
        level < GD_MAX_LEVEL ? info->levelOffsets[level + 1] + info->levelCount[level + 1] : 0;
    const int shift = 63 - 3 * level;
    const unsigned lane = gd_lane();
    const unsigned long long lt = gd_lanemask_lt();
    const int wave = (int)(threadIdx.x >> 6);
    int* sQueue = sQueueAll[wave];
    int4* sRec = sRecAll[wave];
    const int waves = gridDim.x * 4;
    int queued = 0;   /* wave-uniform */
    int zeros = 0;
    for (int c0 = (blockIdx.x * 4 + wave) * 64; c0 < cnt; c0 += waves * 64) {
        const int t = c0 + (int)lane;
        bool inner = false;
        if (t < cnt) {
            const int idx = off + t;
            const int4 mt = meta[idx];
            inner = mt.w >= 0;
            if (!inner) {
                const int first = mt.x, count = mt.y & 0x1fffffff;
                const uint64_t key = keys[first];
                float sx = 0.f, sy = 0.f, sz = 0.f, m = 0.f;
                for (int b = 0; b < count; ++b) {
                    const float4 p = pos[first + b];
                    sx = __builtin_fmaf(p.x, p.w, sx);
                    sy = __builtin_fmaf(p.y, p.w, sy);
                    sz = __builtin_fmaf(p.z, p.w, sz);
                    m += p.w;
                }
                float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
                if (m > 0.f) c = make_float4(sx / m, sy / m, sz / m, m);
                else zeros++;
                {   /* the record, staged in LDS (written out below) */
                    const uint64_t path = key >> shift;
                    int4* r = sRec + 5 * lane;
                    r[0] = make_int4((int)(uint32_t)path, (int)(uint32_t)(path >> 32), mt.z, first);
                    r[1] = make_int4(count, -1, -1, -1);
                    r[2] = make_int4(-1, -1, -1, -1);
                    r[3] = make_int4(-1, level, 0, 0);
                    r[4] = make_int4(__float_as_int(c.x), __float_as_int(c.y), __float_as_int(c.z),
                                     __float_as_int(c.w));
                }
                wcom[idx] = c;
                wlink[idx] = make_int2(-(first + 1), (count << 5) | level);
                if (count > 1) {   /* moments of a single-body leaf are zero, not stored */
                    if (MOM >= 1) {
                        float4 q0, q1;
                        gd_leaf_moments(pos, first, count, c, &q0, &q1);
                        wquad[2 * idx] = q0;
                        wquad[2 * idx + 1] = q1;
                    }
                    if (MOM >= 2) {
                        float4 o0, o1, o2;
                        gd_leaf_third(pos, first, count, c, &o0, &o1, &o2);
                        woct[3 * idx] = o0;
                        woct[3 * idx + 1] = o1;
                        woct[3 * idx + 2] = o2;
                    }
                }
            }
        }
        /* the chunk's leaf records, as 16-byte pieces in address order: lane
           l of store u writes piece 64 u + l of the chunk's 64 records, so a
           store covers 1 KB of consecutive records (less the internal nodes'
           records, written when their node is finished) instead of 64
           pieces 80 B apart */
        {
            const unsigned long long leafMask = __ballot(t < cnt && !inner);
            __builtin_amdgcn_wave_barrier();
            int4* base = reinterpret_cast<int4*>(nodes + off + c0);
#pragma unroll
            for (int u = 0; u < 5; ++u) {
                const int e = 64 * u + (int)lane;
                const int rr = e / 5;
                if ((leafMask >> rr) & 1ull) base[e] = sRec[e];
            }
            __builtin_amdgcn_wave_barrier();
        }
        const unsigned long long bal = __ballot(inner);
        if (inner) sQueue[queued + __popcll(bal & lt)] = off + t;
        queued += __popcll(bal);
        __builtin_amdgcn_wave_barrier();
        if (queued >= 64) {
            const int idx = sQueue[lane];
            const int rest = queued - 64 > (int)lane ? sQueue[64 + lane] : 0;
            __builtin_amdgcn_wave_barrier();
            if (queued - 64 > (int)lane) sQueue[lane] = rest;
            queued -= 64;
            zeros += gd_finish_inner<MOM>(nodes, keys, meta, idx, level, childEnd, wcom, wlink,
                                          wquad, woct);
            __builtin_amdgcn_wave_barrier();
        }
    }
    if ((int)lane < queued)
        zeros += gd_finish_inner<MOM>(nodes, keys, meta, sQueue[lane], level, childEnd, wcom,
                                      wlink, wquad, woct);
    if (zeros) atomicAdd(&info->zeroMassNodes, zeros);
}

hipEvent_t gd_build_mid_event = nullptr;

void gd_build_octree(gd_builder* b, const uint64_t* sortedKeys, const float4* sortedPos, int n,
                     node* nodes, int maxNodes) {
    if (n <= 0) return;
    const int tiles = gd_div_up(n, GD_BUILD_TILE);
    if (tiles > b->maxTiles) {
        GD_FAIL("octree build of %d bodies exceeds the builder's capacity", n);
        return;
    }
    k_build_count<<<tiles, GD_BUILD_THREADS, 0, gd_stream()>>>(sortedKeys, n, b->tileCounts, b->maxTiles);
    GD_LAUNCHED();
    k_build_scan<<<GD_MAX_LEVEL, 1024, 0, gd_stream()>>>(b->tileCounts, b->tileOffsets, tiles, b->maxTiles,
                                         b->levelTotals);
    GD_LAUNCHED();
    k_build_nodes<<<tiles, GD_BUILD_THREADS, 0, gd_stream()>>>(sortedKeys, n, b->tileOffsets, b->levelTotals,
                                               b->maxTiles, b->meta, maxNodes, b->dinfo);
    GD_LAUNCHED();
    if (gd_build_mid_event) GD_HIP(hipEventRecord(gd_build_mid_event, gd_stream()));
    for (int L = GD_MAX_LEVEL; L >= 0; --L) {
        /* a level holds at most min(8^L, n) nodes */
        long long bound = n;
        if (L < 8) {
            const long long p8 = 1ll << (3 * L);
            bound = p8 < bound ? p8 : bound;
        }
        int blocks = gd_div_up(bound, 256);
        if (blocks > 4096) blocks = 4096;   /* 16 blocks per CU */
        if (b->quad == 2)
            k_build_mass<2><<<blocks, 256, 0, gd_stream()>>>(nodes, sortedPos, sortedKeys, b->meta, b->dinfo, L,
                                             b->wcom, b->wlink, b->wquad, b->woct);
        else if (b->quad == 1)
            k_build_mass<1><<<blocks, 256, 0, gd_stream()>>>(nodes, sortedPos, sortedKeys, b->meta, b->dinfo, L,
                                             b->wcom, b->wlink, b->wquad, nullptr);
        else
            k_build_mass<0><<<blocks, 256, 0, gd_stream()>>>(nodes, sortedPos, sortedKeys, b->meta, b->dinfo, L,
                                             b->wcom, b->wlink, nullptr, nullptr);
        GD_LAUNCHED();
    }
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
