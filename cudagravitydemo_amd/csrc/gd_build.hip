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
 *                   (equal to tree.cu:181-203's sibling / parent-end rule);
 *                   child[] = -1
 *   k_build_mass    one launch per level, deepest first (tree.cu:89-179, same
 *                   operation order as the oracle); each node also writes its
 *                   own slot in the parent's child[] (tree.cu:232-233), and its
 *                   compact walk record (centre of mass, first child / body,
 *                   count, level)
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
    const uint32_t* __restrict__ levelTotals, int maxTiles, node* __restrict__ nodes,
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
            if (!overflow) {
                /* root, basicGravity.cu:158-169 */
                int4* r = reinterpret_cast<int4*>(nodes);
                r[0] = make_int4(0, 0, -1, 0);
                r[1] = make_int4(n, -1, -1, -1);
                r[2] = make_int4(-1, -1, -1, -1);
                r[3] = make_int4(-1, 0, 0, 0);
            }
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
        for (int L = 1; L <= waveMaxHi; ++L) {
            const bool flag = lo[k] <= L && L <= hi[k];
            const unsigned long long bal = __ballot(flag);
            const int rank = (int)sRank[k][wave][L] + __popcll(bal & lt);
            if (flag) {
                const int idx = sLevelOff[L] + rank;
                const int parent = (L == 1) ? 0 : sLevelOff[L - 1] + prevRank + prevFlag - 1;
                const int shift = 63 - 3 * L;
                const uint64_t cell = key >> shift;
                const int count = (L > Bv[k]) ? 1 : gd_cell_size(keys, n, i, cell, shift);
                int4* w = reinterpret_cast<int4*>(nodes + idx);
                w[0] = make_int4((int)(uint32_t)cell, (int)(uint32_t)(cell >> 32), parent, (int)i);
                w[1] = make_int4(count, -1, -1, -1);
                w[2] = make_int4(-1, -1, -1, -1);
                w[3] = make_int4(-1, L, 0, 0);
            }
            prevRank = rank;
            prevFlag = flag;
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

/* Second moments about the centre of mass, S = sum m (x - c)(x - c)^T, as
 * two float4: (Sxx, Syy, Szz, Sxy), (Sxz, Syz, trace, 0).  Leaves sum over
 * their bodies; internal nodes shift each child's moments by the parallel-axis
 * term m_c (c_c - c)(c_c - c)^T. */
__device__ __forceinline__ void gd_node_moments(const float4* __restrict__ pos,
                                                const float4* __restrict__ wcom,
                                                const int2* __restrict__ wlink,
                                                const float4* __restrict__ wquad, int first,
                                                int count, const int* child, bool leaf, float4 c,
                                                float4* q0, float4* q1) {
    float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
    if (leaf) {
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
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (child[k] == -1) continue;
            const float4 cc = wcom[child[k]];
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
            if (!gd_single_leaf(wlink, child[k])) {
                a = wquad[2 * child[k]];
                b = wquad[2 * child[k] + 1];
            }
            const float dx = cc.x - c.x, dy = cc.y - c.y, dz = cc.z - c.z;
            sxx += __builtin_fmaf(cc.w * dx, dx, a.x);
            syy += __builtin_fmaf(cc.w * dy, dy, a.y);
            szz += __builtin_fmaf(cc.w * dz, dz, a.z);
            sxy += __builtin_fmaf(cc.w * dx, dy, a.w);
            sxz += __builtin_fmaf(cc.w * dx, dz, b.x);
            syz += __builtin_fmaf(cc.w * dy, dz, b.y);
        }
    }
    *q0 = make_float4(sxx, syy, szz, sxy);
    *q1 = make_float4(sxz, syz, sxx + syy + szz, 0.f);
}

/* Third moments O_ijk = sum m d_i d_j d_k about the COM c (10 components:
 * xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz).  Leaves sum over their bodies;
 * internal nodes shift each child's moments to c (d = c_child - c):
 * O += O_c + (S_c,ij d_k + S_c,ik d_j + S_c,jk d_i) + m_c d_i d_j d_k. */
__device__ __forceinline__ void gd_node_third(const float4* __restrict__ pos,
                                              const float4* __restrict__ wcom,
                                              const int2* __restrict__ wlink,
                                              const float4* __restrict__ wquad,
                                              const float4* __restrict__ woct, int first,
                                              int count, const int* child, bool leaf, float4 c,
                                              float4* o0, float4* o1, float4* o2) {
    float xxx = 0.f, yyy = 0.f, zzz = 0.f, xxy = 0.f, xxz = 0.f;
    float xyy = 0.f, yyz = 0.f, xzz = 0.f, yzz = 0.f, xyz = 0.f;
    if (leaf) {
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
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (child[k] == -1) continue;
            const int j = child[k];
            const float4 cc = wcom[j];
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, p = a, q = a, r = a;
            if (!gd_single_leaf(wlink, j)) {
                a = wquad[2 * j];
                b = wquad[2 * j + 1];
                p = woct[3 * j];
                q = woct[3 * j + 1];
                r = woct[3 * j + 2];
            }
            const float sxx = a.x, syy = a.y, szz = a.z, sxy = a.w, sxz = b.x, syz = b.y;
            const float dx = cc.x - c.x, dy = cc.y - c.y, dz = cc.z - c.z;
            const float mx = cc.w * dx, my = cc.w * dy, mz = cc.w * dz;
            xxx += p.x + 3.f * sxx * dx + mx * dx * dx;
            yyy += p.y + 3.f * syy * dy + my * dy * dy;
            zzz += p.z + 3.f * szz * dz + mz * dz * dz;
            xxy += p.w + sxx * dy + 2.f * sxy * dx + mx * dx * dy;
            xxz += q.x + sxx * dz + 2.f * sxz * dx + mx * dx * dz;
            xyy += q.y + syy * dx + 2.f * sxy * dy + mx * dy * dy;
            yyz += q.z + syy * dz + 2.f * syz * dy + my * dy * dz;
            xzz += q.w + szz * dx + 2.f * sxz * dz + mx * dz * dz;
            yzz += r.x + szz * dy + 2.f * syz * dz + my * dz * dz;
            xyz += r.y + sxy * dz + sxz * dy + syz * dx + mx * dy * dz;
        }
    }
    *o0 = make_float4(xxx, yyy, zzz, xxy);
    *o1 = make_float4(xxz, xyy, yyz, xzz);
    *o2 = make_float4(yzz, xyz, 0.f, 0.f);
}

/* MOM: 0 centre of mass only, 1 also second moments, 2 also third moments */
template <int MOM>
__global__ __launch_bounds__(256) void k_build_mass(node* __restrict__ nodes,
                                                    const float4* __restrict__ pos,
                                                    gd_tree_info* info, int level,
                                                    float4* __restrict__ wcom,
                                                    int2* __restrict__ wlink,
                                                    float4* __restrict__ wquad,
                                                    float4* __restrict__ woct) {
    if (info->overflow) return;
    const int cnt = info->levelCount[level];
    const int off = info->levelOffsets[level];
    int zeros = 0;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < cnt; t += gridDim.x * 256) {
        const int idx = off + t;
        const int4* r = reinterpret_cast<const int4*>(nodes + idx);
        const int4 w0 = r[0], w1 = r[1], w2 = r[2], w3 = r[3];
        const int child[8] = {w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w, w3.x};
        const int first = w0.w, count = w1.x;
        const gd_mass_result m =
            gd_node_mass([wcom](int j) { return wcom[j]; }, pos, first, count, child);
        nodes[idx].massData = m.m;
        zeros += m.zero;
        if (level > 0) {
            const int parent = w0.z;
            nodes[parent].child[(uint32_t)w0.x & 7u] = idx;
        }
        wcom[idx] = m.m;
        wlink[idx] = m.leaf ? make_int2(-(first + 1), (count << 5) | level)
                            : make_int2(m.firstChild, (m.nChild << 5) | level);
        const bool single = m.leaf && count == 1;   /* moments zero, not stored */
        if (MOM >= 1 && !single) {
            float4 q0, q1;
            gd_node_moments(pos, wcom, wlink, wquad, first, count, child, m.leaf, m.m, &q0, &q1);
            wquad[2 * idx] = q0;
            wquad[2 * idx + 1] = q1;
        }
        if (MOM >= 2 && !single) {
            float4 o0, o1, o2;
            gd_node_third(pos, wcom, wlink, wquad, woct, first, count, child, m.leaf, m.m, &o0, &o1,
                          &o2);
            woct[3 * idx] = o0;
            woct[3 * idx + 1] = o1;
            woct[3 * idx + 2] = o2;
        }
    }
    if (zeros) atomicAdd(&info->zeroMassNodes, zeros);
}

void gd_build_octree(gd_builder* b, const uint64_t* sortedKeys, const float4* sortedPos, int n,
                     node* nodes, int maxNodes) {
    if (n <= 0) return;
    const int tiles = gd_div_up(n, GD_BUILD_TILE);
    if (tiles > b->maxTiles) {
        GD_FAIL("octree build of %d bodies exceeds the builder's capacity", n);
        return;
    }
    k_build_count<<<tiles, GD_BUILD_THREADS>>>(sortedKeys, n, b->tileCounts, b->maxTiles);
    GD_LAUNCHED();
    k_build_scan<<<GD_MAX_LEVEL, 1024>>>(b->tileCounts, b->tileOffsets, tiles, b->maxTiles,
                                         b->levelTotals);
    GD_LAUNCHED();
    k_build_nodes<<<tiles, GD_BUILD_THREADS>>>(sortedKeys, n, b->tileOffsets, b->levelTotals,
                                               b->maxTiles, nodes, maxNodes, b->dinfo);
    GD_LAUNCHED();
    for (int L = GD_MAX_LEVEL; L >= 0; --L) {
        /* a level holds at most min(8^L, n) nodes */
        long long bound = n;
        if (L < 8) {
            const long long p8 = 1ll << (3 * L);
            bound = p8 < bound ? p8 : bound;
        }
        int blocks = gd_div_up(bound, 256);
        if (blocks > 4096) blocks = 4096;   /* 16 blocks per CU: every wave slot busy */
        if (b->quad == 2)
            k_build_mass<2><<<blocks, 256>>>(nodes, sortedPos, b->dinfo, L, b->wcom, b->wlink,
                                             b->wquad, b->woct);
        else if (b->quad == 1)
            k_build_mass<1><<<blocks, 256>>>(nodes, sortedPos, b->dinfo, L, b->wcom, b->wlink,
                                             b->wquad, nullptr);
        else
            k_build_mass<0><<<blocks, 256>>>(nodes, sortedPos, b->dinfo, L, b->wcom, b->wlink,
                                             nullptr, nullptr);
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
    k_identify<<<gd_div_up(starCount, 256), 256>>>(mortonCodes, nodeLevel, starCount, nodeFlags);
    GD_LAUNCHED();
}

void createNodes(const uint64_t* mortonCodes, const uint32_t* flags, const uint32_t* previousFlags,
                 const uint32_t* offsets, const uint32_t* offsetsPrev, node* nodes, int level,
                 int arrayLevelOffset, int prevArrayLevelOffset, int starCount, int maxNodes) {
    if (starCount <= 0) return;
    k_create_nodes<<<gd_div_up(starCount, 256), 256>>>(mortonCodes, flags, previousFlags, offsets,
                                                       offsetsPrev, nodes, level, arrayLevelOffset,
                                                       prevArrayLevelOffset, starCount, maxNodes);
    GD_LAUNCHED();
}

void updateTreeParticles(node* nodes, int levelStartIndex, int levelNodeCount, int starCount) {
    (void)starCount;
    if (levelNodeCount <= 0) return;
    k_tree_particles<<<gd_div_up(levelNodeCount, 256), 256>>>(nodes, levelStartIndex,
                                                               levelNodeCount);
    GD_LAUNCHED();
}

void updateTreeMass(node* nodes, const float4* positionMassVals, int level, int levelNodes,
                    int levelStartIndex) {
    (void)level;
    if (levelNodes <= 0) return;
    k_tree_mass_api<<<gd_div_up(levelNodes, 256), 256>>>(nodes, positionMassVals, levelNodes,
                                                         levelStartIndex);
    GD_LAUNCHED();
}

}  /* extern "C" */
