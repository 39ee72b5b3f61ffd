/*
 * gd_sort.hip -- stable LSD radix sort of (u64 key, u32 value) pairs, and the
 * u32 exclusive scan behind prefixSum.
 *
 * Replaces the reference's CUB calls: DeviceRadixSort::SortPairs over bits
 * [0,64) (sorting.cu:41,77) and DeviceScan::ExclusiveSum (tree.cu:255,297).
 * Semantics: the permutation equals a stable sort by key; offsets are the
 * exclusive running sum.
 *
 * Sort design (gfx950, 64-wide wavefronts):
 *   k_sort_init    clears the per-range digit histograms and tile tickets
 *   k_sort_hist    before each pass, that pass's digit counts per tile range
 *                  (below); each thread walks 8 consecutive keys and adds runs
 *                  of equal digits at once, first into LDS, then into global
 *                  counters
 *   k_sort_prefix  that pass's counts -> each range's digit offsets
 *   k_sort_pass x8 one 8-bit digit per pass over 3840-key tiles:
 *                  - the tiles are split into GD_SORT_GROUPS (64) contiguous
 *                    ranges, eight per XCD (block b runs on XCD b % 8, range
 *                    g on XCD g % 8): the blocks b = g mod 64 take the tiles
 *                    of range g in order from an atomic ticket of their own,
 *                    so a tile only ever waits on tiles of its range that
 *                    started before it (64 short look-back chains instead of
 *                    8 long ones: 0.137 -> 0.118 ms per pass at 2^24 keys;
 *                    16 and 32 ranges 0.132 and 0.123, 128 no better than 64),
 *                    and consecutive tiles -- which write
 *                    the two halves of the 64-byte sectors where their runs
 *                    of a digit meet -- run in one L2, where those halves
 *                    merge instead of each costing the memory a
 *                    read-modify-write (a digit's run in a tile is ~15 keys,
 *                    so most sectors are such boundaries);
 *                  - the histogram pass counts each range's digits apart, so
 *                    each range starts from its own offsets and the
 *                    look-back never leaves its range;
 *                  - keys are ranked inside the tile with wavefront ballots
 *                    (8 ballots give the lanes sharing a digit) and per-wave
 *                    digit counters in LDS, which keeps the order stable;
 *                  - per-digit tile counts are published in 64-bit status
 *                    words {epoch, flag, count} and each tile looks back over
 *                    its predecessors (decoupled look-back) for its global
 *                    offset; the epoch tag makes words of earlier passes
 *                    read as "not ready", so the status array is never
 *                    cleared between passes;
 *                  - keys are staged in LDS in digit order and written out
 *                    as runs of consecutive addresses.
 * Traffic per key: 8 x (8 B histogram read + 12 B read + 12 B write).
 */
#include "gd_internal.h"

#define GD_FLAG_AGG 1u
#define GD_FLAG_PREFIX 2u

/* ---------------------------------------------------------------------------
 * Block scan helpers
 * ------------------------------------------------------------------------- */

/* Inclusive scan across one wavefront. */
__device__ __forceinline__ uint32_t gd_wave_inclusive(uint32_t v) {
    const unsigned lane = gd_lane();
#pragma unroll
    for (int off = 1; off < GD_WAVE; off <<= 1) {
        const uint32_t o = __shfl_up(v, off);
        if (lane >= (unsigned)off) v += o;
    }
    return v;
}

/* Exclusive scan across a block of THREADS threads; *total gets the sum. */
template <int THREADS>
__device__ __forceinline__ uint32_t gd_block_exclusive(uint32_t v, uint32_t* total) {
    __shared__ uint32_t sWave[THREADS / GD_WAVE];
    __shared__ uint32_t sTotal;
    const int wave = threadIdx.x / GD_WAVE;
    const uint32_t inc = gd_wave_inclusive(v);
    if (gd_lane() == GD_WAVE - 1) sWave[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < THREADS / GD_WAVE; ++w) {
            const uint32_t c = sWave[w];
            sWave[w] = run;
            run += c;
        }
        sTotal = run;
    }
    __syncthreads();
    const uint32_t out = sWave[wave] + inc - v;
    if (total) *total = sTotal;
    __syncthreads();
    return out;
}

/* ---------------------------------------------------------------------------
 * Radix sort kernels
 * ------------------------------------------------------------------------- */
/* Clears the digit counts and tile tickets and advances the status-word
 * epoch by one per pass, on the device (so a captured sort replays with new
 * tags): passes p = 0..7 tag their words *epoch + p; tags stay in
 * [1, 2^30 - 2], 0 being the tag of a cleared word. */
__global__ void k_sort_init(uint32_t* hist, uint32_t* tileCounter, uint32_t* epoch) {
    const int t = threadIdx.x;   /* 256 threads */
    for (int p = 0; p < GD_SORT_PASSES * GD_SORT_GROUPS; ++p) hist[p * GD_SORT_BINS + t] = 0u;
    for (int k = t; k < GD_SORT_PASSES * GD_SORT_GROUPS; k += 256) tileCounter[k] = 0u;
    if (t == 0) {
        uint32_t e = *epoch + GD_SORT_PASSES;
        if (e == GD_SORT_PASSES || e > 0x3fffffffu - GD_SORT_PASSES) e = 1u;
        *epoch = e;
    }
}

#define GD_HIST_ITEMS 8
#define GD_HIST_CHUNK (256 * GD_HIST_ITEMS)

/* Range g of the tiles: the blocks b = g, g + G, ... of a grid of one block
 * per tile (G = GD_SORT_GROUPS), i.e. (tiles - g + G - 1) / G tiles, the
 * first tiles % G ranges one more than the others, in order. */
__device__ __forceinline__ void gd_sort_range(int tiles, int g, int* start, int* count) {
    const int q = tiles / GD_SORT_GROUPS, r = tiles % GD_SORT_GROUPS;
    *start = g * q + (g < r ? g : r);
    *count = q + (g < r ? 1 : 0);
}

/* hist[g][d] (range g's count of digit d) -> the global position where
 * range g's keys of digit d start: all keys of smaller digits, then those
 * of digit d in earlier ranges.  One block of 256 threads, thread = digit. */
__device__ __forceinline__ void gd_sort_offsets(uint32_t* h) {
    const int d = threadIdx.x;
    uint32_t c[GD_SORT_GROUPS], tot = 0;
#pragma unroll
    for (int g = 0; g < GD_SORT_GROUPS; ++g) {
        c[g] = h[g * GD_SORT_BINS + d];
        tot += c[g];
    }
    uint32_t run = gd_block_exclusive<256>(tot, nullptr);
#pragma unroll
    for (int g = 0; g < GD_SORT_GROUPS; ++g) {
        h[g * GD_SORT_BINS + d] = run;
        run += c[g];
    }
}

/* One pass's digit counts per tile range, over that pass's input: blocks
 * g, g + 8, ... count range g's keys into hist[g][digit] (the ranges'
 * counts change from pass to pass, as the keys move between ranges). */
__global__ __launch_bounds__(256) void k_sort_hist(const uint64_t* __restrict__ keys, int n,
                                                   int shift, uint32_t* hist) {
    __shared__ uint32_t sh[GD_SORT_BINS];
    const int t = threadIdx.x;
    const int g = blockIdx.x % GD_SORT_GROUPS;
    const int nb = (gridDim.x - g + GD_SORT_GROUPS - 1) / GD_SORT_GROUPS;   /* blocks of range g */
    const int bi = blockIdx.x / GD_SORT_GROUPS;
    int rs, rc;
    gd_sort_range((int)(((long long)n + GD_SORT_TILE - 1) / GD_SORT_TILE), g, &rs, &rc);
    const long long k0 = (long long)rs * GD_SORT_TILE;
    const long long kEnd = (long long)(rs + rc) * GD_SORT_TILE;
    const long long k1 = kEnd < n ? kEnd : n;
    sh[t] = 0u;
    __syncthreads();
    for (long long base = k0 + (long long)bi * GD_HIST_CHUNK; base < k1;
         base += (long long)nb * GD_HIST_CHUNK) {
        const long long i0 = base + (long long)t * GD_HIST_ITEMS;
        if (i0 >= k1) continue;
        const int cnt = (k1 - i0) < GD_HIST_ITEMS ? (int)(k1 - i0) : GD_HIST_ITEMS;
        uint64_t k[GD_HIST_ITEMS];
        if (cnt == GD_HIST_ITEMS) {
            const ulonglong2* src = reinterpret_cast<const ulonglong2*>(keys + i0);
#pragma unroll
            for (int j = 0; j < GD_HIST_ITEMS / 2; ++j) {
                const ulonglong2 v = src[j];
                k[2 * j] = v.x;
                k[2 * j + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < GD_HIST_ITEMS; ++j) k[j] = j < cnt ? keys[i0 + j] : 0ull;
        }
        uint32_t prev = (uint32_t)(k[0] >> shift) & 0xffu;
        uint32_t run = 1;
        for (int j = 1; j < cnt; ++j) {
            const uint32_t d = (uint32_t)(k[j] >> shift) & 0xffu;
            if (d == prev) {
                ++run;
            } else {
                atomicAdd(&sh[prev], run);
                prev = d;
                run = 1;
            }
        }
        atomicAdd(&sh[prev], run);
    }
    __syncthreads();
    const uint32_t c = sh[t];
    if (c) atomicAdd(&hist[g * GD_SORT_BINS + t], c);
}

/* One pass's counts -> offsets (one block of 256 threads). */
__global__ __launch_bounds__(256) void k_sort_prefix(uint32_t* hist) { gd_sort_offsets(hist); }

__device__ __forceinline__ unsigned long long gd_pack_status(uint32_t epoch, uint32_t flag,
                                                             uint32_t count) {
    return ((unsigned long long)(((epoch & 0x3fffffffu) << 2) | flag) << 32) | count;
}

__global__ __launch_bounds__(GD_SORT_THREADS) void k_sort_pass(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, int n, int shift,
    const uint32_t* __restrict__ digitBase, uint32_t* tileCounter,
    unsigned long long* status, const uint32_t* __restrict__ epochBase, int pass) {
    constexpr int WAVES = GD_SORT_THREADS / GD_WAVE;
    constexpr int WAVE_SPAN = GD_SORT_ITEMS * GD_WAVE;   /* keys per wave */
    __shared__ uint64_t sKeys[GD_SORT_TILE];
    __shared__ uint32_t sVals[GD_SORT_TILE];
    __shared__ uint32_t sWaveHist[WAVES][GD_SORT_BINS];
    __shared__ uint32_t sTileStart[GD_SORT_BINS];
    __shared__ uint32_t sGlobal[GD_SORT_BINS];
    __shared__ uint32_t sTile;

    const int t = threadIdx.x;
    const unsigned lane = gd_lane();
    const int wave = t / GD_WAVE;
    const uint32_t epoch = *epochBase + (uint32_t)pass;
    const int g = blockIdx.x % GD_SORT_GROUPS;   /* the tile range (on XCD g % 8) */
    if (t == 0) sTile = atomicAdd(&tileCounter[g], 1u);
#pragma unroll
    for (int w = 0; w < WAVES; ++w) sWaveHist[w][t] = 0u;
    __syncthreads();
    int rs, rc;
    gd_sort_range((int)gridDim.x, g, &rs, &rc);
    const int inRange = (int)sTile;               /* < rc: range g has rc blocks */
    const int tile = rs + inRange;
    const long long tileBase = (long long)tile * GD_SORT_TILE;
    digitBase += g * GD_SORT_BINS;

    /* load: item j of this lane sits at tileBase + wave*WAVE_SPAN + j*64 + lane */
    uint64_t key[GD_SORT_ITEMS];
    uint32_t val[GD_SORT_ITEMS];
    uint32_t packed[GD_SORT_ITEMS];   /* digit << 16 | rank inside the wave */
#pragma unroll
    for (int j = 0; j < GD_SORT_ITEMS; ++j) {
        const long long p = tileBase + wave * WAVE_SPAN + j * GD_WAVE + lane;
        if (p < n) {
            key[j] = kin[p];
            val[j] = vin[p];
        } else {
            key[j] = 0ull;
            val[j] = 0u;
        }
    }
    const unsigned long long lt = gd_lanemask_lt();
#pragma unroll
    for (int j = 0; j < GD_SORT_ITEMS; ++j) {
        const long long p = tileBase + wave * WAVE_SPAN + j * GD_WAVE + lane;
        const bool valid = p < n;
        const uint32_t d = (uint32_t)(key[j] >> shift) & 0xffu;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t rank = 0;
        if (valid) {
            const uint32_t before = (uint32_t)__popcll(peers & lt);
            const uint32_t prior = sWaveHist[wave][d];
            rank = prior + before;
            if (before == 0) sWaveHist[wave][d] = prior + (uint32_t)__popcll(peers);
        }
        packed[j] = (d << 16) | rank;
        /* every lane's read of this round precedes the next round's update */
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    /* per digit (thread t owns digit t): counts of the waves, tile total */
    uint32_t tot = 0;
    {
        uint32_t c[WAVES];
#pragma unroll
        for (int w = 0; w < WAVES; ++w) c[w] = sWaveHist[w][t];
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            sWaveHist[w][t] = tot;
            tot += c[w];
        }
    }
    unsigned long long* myStatus = status + (size_t)tile * GD_SORT_BINS + t;
    if (inRange == 0) {
        __hip_atomic_store(myStatus, gd_pack_status(epoch, GD_FLAG_PREFIX, tot),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(myStatus, gd_pack_status(epoch, GD_FLAG_AGG, tot), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint32_t tileStart = gd_block_exclusive<GD_SORT_THREADS>(tot, nullptr);
    sTileStart[t] = tileStart;

    /* decoupled look-back for digit t (the range's first tile publishes an
       inclusive prefix, so the walk never leaves the range) */
    uint32_t excl = 0;
    if (inRange > 0) {
        const uint32_t want = (epoch & 0x3fffffffu);
        int pred = tile - 1;
        while (true) {
            const unsigned long long s =
                __hip_atomic_load(status + (size_t)pred * GD_SORT_BINS + t, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t hi = (uint32_t)(s >> 32);
            const uint32_t flag = hi & 3u;
            if ((hi >> 2) != want || flag == 0u) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += (uint32_t)s;
            if (flag == GD_FLAG_PREFIX) break;
            --pred;
        }
        __hip_atomic_store(myStatus, gd_pack_status(epoch, GD_FLAG_PREFIX, excl + tot),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    sGlobal[t] = digitBase[t] + excl - tileStart;
    __syncthreads();

    /* stage the tile in digit order */
#pragma unroll
    for (int j = 0; j < GD_SORT_ITEMS; ++j) {
        const long long p = tileBase + wave * WAVE_SPAN + j * GD_WAVE + lane;
        if (p < n) {
            const uint32_t d = packed[j] >> 16;
            const uint32_t local = sTileStart[d] + sWaveHist[wave][d] + (packed[j] & 0xffffu);
            sKeys[local] = key[j];
            sVals[local] = val[j];
        }
    }
    __syncthreads();

    const int count = (n - tileBase) < GD_SORT_TILE ? (int)(n - tileBase) : GD_SORT_TILE;
    for (int q = t; q < count; q += GD_SORT_THREADS) {
        const uint64_t k = sKeys[q];
        const uint32_t d = (uint32_t)(k >> shift) & 0xffu;
        const uint32_t g = sGlobal[d] + (uint32_t)q;
        kout[g] = k;
        vout[g] = sVals[q];
    }
}

void gd_sort_pairs(gd_sorter* s, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                   uint32_t* vout, int n) {
    if (n <= 0) return;
    const int tiles = gd_div_up(n, GD_SORT_TILE);
    if (tiles > s->maxTiles) {
        GD_FAIL("radix sort of %d keys exceeds the sorter's %d tiles", n, s->maxTiles);
        return;
    }
    k_sort_init<<<1, 256, 0, gd_stream()>>>(s->hist, s->tileCounter, s->epoch);
    GD_LAUNCHED();
    int hb = gd_div_up(n, GD_HIST_CHUNK);
    if (hb > 1024) hb = 1024;
    hb = gd_div_up(hb, GD_SORT_GROUPS) * GD_SORT_GROUPS;   /* whole rounds of the ranges */
    const uint64_t* ksrc = kin;
    const uint32_t* vsrc = vin;
    for (int p = 0; p < GD_SORT_PASSES; ++p) {
        uint32_t* h = s->hist + p * GD_SORT_GROUPS * GD_SORT_BINS;
        k_sort_hist<<<hb, 256, 0, gd_stream()>>>(ksrc, n, 8 * p, h);
        GD_LAUNCHED();
        k_sort_prefix<<<1, 256, 0, gd_stream()>>>(h);
        GD_LAUNCHED();
        /* even passes write the scratch pair, odd passes the output pair, so
           the eighth pass lands in keysOut/valsOut */
        uint64_t* kdst = (p & 1) ? kout : s->keysTmp;
        uint32_t* vdst = (p & 1) ? vout : s->valsTmp;
        k_sort_pass<<<tiles, GD_SORT_THREADS, 0, gd_stream()>>>(ksrc, vsrc, kdst, vdst, n, 8 * p,
                                                s->hist + p * GD_SORT_GROUPS * GD_SORT_BINS,
                                                s->tileCounter + p * GD_SORT_GROUPS,
                                                s->status, s->epoch, p);
        GD_LAUNCHED();
        ksrc = kdst;
        vsrc = vdst;
    }
}

/* ---------------------------------------------------------------------------
 * Exclusive scan of u32 (prefixSum): tile sums, scan of the sums, tiles.
 * ------------------------------------------------------------------------- */
__global__ __launch_bounds__(GD_SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                                 int n, uint32_t* tileSums) {
    const long long base = (long long)blockIdx.x * GD_SCAN_TILE + threadIdx.x * GD_SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < GD_SCAN_ITEMS; ++j)
        if (base + j < n) s += in[base + j];
    uint32_t total;
    gd_block_exclusive<GD_SCAN_THREADS>(s, &total);
    if (threadIdx.x == 0) tileSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* tileSums, int tiles) {
    __shared__ uint32_t sCarry;
    if (threadIdx.x == 0) sCarry = 0;
    __syncthreads();
    for (int base = 0; base < tiles; base += 1024) {
        const int i = base + threadIdx.x;
        const uint32_t v = i < tiles ? tileSums[i] : 0u;
        uint32_t total;
        const uint32_t ex = gd_block_exclusive<1024>(v, &total);
        const uint32_t carry = sCarry;
        if (i < tiles) tileSums[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) sCarry = carry + total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(GD_SCAN_THREADS) void k_scan_down(const uint32_t* __restrict__ in,
                                                               uint32_t* __restrict__ out, int n,
                                                               const uint32_t* tileSums) {
    const long long base = (long long)blockIdx.x * GD_SCAN_TILE + threadIdx.x * GD_SCAN_ITEMS;
    uint32_t v[GD_SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < GD_SCAN_ITEMS; ++j) {
        v[j] = base + j < n ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t run = gd_block_exclusive<GD_SCAN_THREADS>(s, nullptr) + tileSums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < GD_SCAN_ITEMS; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
}

void gd_exclusive_scan_u32(gd_builder* b, const uint32_t* in, uint32_t* out, int n) {
    if (n <= 0) return;
    const int tiles = gd_div_up(n, GD_SCAN_TILE);
    k_scan_reduce<<<tiles, GD_SCAN_THREADS, 0, gd_stream()>>>(in, n, b->scanTiles);
    GD_LAUNCHED();
    k_scan_sums<<<1, 1024, 0, gd_stream()>>>(b->scanTiles, tiles);
    GD_LAUNCHED();
    k_scan_down<<<tiles, GD_SCAN_THREADS, 0, gd_stream()>>>(in, out, n, b->scanTiles);
    GD_LAUNCHED();
}
