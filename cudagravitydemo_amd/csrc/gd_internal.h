/*
 * gd_internal.h -- shared by the engine's HIP translation units.
 *
 * Error convention: like the reference's CUDA_CHECK (macros.h:8-17), an error
 * prints where it happened and exits with -1.  Tests switch that off with
 * gd_set_exit_on_error(0); the first error is then kept for gd_last_error()
 * and the failing entry point returns without doing its work.
 */
#ifndef GD_INTERNAL_H
#define GD_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "gravity_types.h"
#include "gravity_ext.h"

/* ---- error handling ---------------------------------------------------- */
void gd_fail(const char* file, int line, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int gd_has_error(void);
/* The volume fraction below which the walks split a 64-body run in two (0:
 * whole runs): gd_walk_set_group_split's setting, else the rule for this
 * opening angle. */
float gd_walk_split_ratio(float theta);

/* The stream of every launch, asynchronous copy and event of the library. */
hipStream_t gd_stream(void);
hipStream_t gd_side_stream(void);
/* 1 while a gd_graph_begin capture is open. */
int gd_capturing(void);
/* 1 while a capture is open or an instantiated graph lives: the buffers a
 * graph's kernels address must then stay where they are. */
int gd_graphs_hold_buffers(void);

#define GD_FAIL(...) gd_fail(__FILE__, __LINE__, __VA_ARGS__)
#define GD_HIP(call)                                                              \
    do {                                                                          \
        hipError_t gd_err_ = (call);                                              \
        if (gd_err_ != hipSuccess)                                                \
            gd_fail(__FILE__, __LINE__, "HIP error %d (%s) in %s", (int)gd_err_,  \
                    hipGetErrorString(gd_err_), #call);                           \
    } while (0)
/* launch check: a bad launch configuration is reported at the call site */
#define GD_LAUNCHED() GD_HIP(hipGetLastError())

/* ---- sizes ------------------------------------------------------------- */
#define GD_WAVE 64
#define GD_MAX_LEVEL 21
#define GD_LEVELS 22                 /* levels 0..21 */

/* bbox: partial min/max blocks of the engine's own reduction */
#define GD_BBOX_BLOCKS 1024
#define GD_BBOX_THREADS 256

/* octree build: one tile = 256 threads x 4 bodies */
#define GD_BUILD_THREADS 256
#define GD_BUILD_ITEMS 4
#define GD_BUILD_TILE (GD_BUILD_THREADS * GD_BUILD_ITEMS)

/* radix sort: one tile = 256 threads x 16 keys, 8-bit digits */
#define GD_SORT_THREADS 256
#define GD_SORT_ITEMS 15
#define GD_SORT_TILE (GD_SORT_THREADS * GD_SORT_ITEMS)
#define GD_SORT_PASSES 8
#define GD_SORT_BINS 256
#define GD_SORT_GROUPS 64           /* tile ranges, eight per XCD */

/* exclusive scan (prefixSum): one tile = 256 threads x 8 items */
#define GD_SCAN_THREADS 256
#define GD_SCAN_ITEMS 8
#define GD_SCAN_TILE (GD_SCAN_THREADS * GD_SCAN_ITEMS)

static inline int gd_div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

/*
 * Device-resident summary of the current tree, written by the build kernels
 * so that no host round trip is needed inside a step.
 */
typedef struct {
    int levelOffsets[GD_LEVELS];
    int levelCount[GD_LEVELS];
    int deepestLevel;
    int totalNodes;
    int overflow;                    /* 1 when the node array was too small */
    int zeroMassNodes;               /* nodes whose mass summed to zero */
    worldBox box;                    /* quantisation box of this step */
    float3 bmin, bmax;               /* global bbox corners */
    double rootCom[4];               /* FP64 sums over all bodies: centre of mass x, y, z
                                        (hi + lo positions) and mass */
} gd_tree_info;

/*
 * The sorter handle the library hands out.  Its first member is the public
 * radixSorter, so a radixSorter* from sorterCreate points at one of these.
 */
typedef struct {
    radixSorter pub;
    uint64_t* keysTmp;               /* ping-pong key buffer, capacity */
    uint32_t* valsTmp;               /* ping-pong value buffer, capacity */
    uint32_t* hist;                  /* [8 passes][8 ranges][256] digit counts, then offsets */
    uint32_t* tileCounter;           /* [8 passes][8 ranges] tile tickets */
    unsigned long long* status;      /* [maxTiles][256] look-back status words */
    int maxTiles;
    uint32_t* epoch;                 /* device: tag of this sort's first pass (passes add 0..7) */
} gd_sorter;

/*
 * The builder handle the library hands out (first member: the public
 * treeBuilder).  Besides the reference's flag/offset arrays it owns the
 * device scratch of the single-pass octree build and the compact walk arrays.
 */
typedef struct {
    treeBuilder pub;
    int maxTiles;                    /* build tiles at capacity */
    uint32_t* tileCounts;            /* [22][maxTiles] node starts per level and tile */
    uint32_t* tileOffsets;           /* [22][maxTiles] exclusive scan of tileCounts */
    int* tileFirst;                  /* [22][maxTiles] first body of the tile that begins a
                                        new cell of the level (n if none) */
    int* tileNext;                   /* [22][maxTiles] the same for the bodies after the tile */
    uint32_t* levelTotals;           /* [22] */
    gd_tree_info* dinfo;             /* device */
    gd_tree_info* hinfo;             /* pinned host mirror */
    float3* bboxPartMin;             /* [GD_BBOX_BLOCKS] */
    float3* bboxPartMax;
    double* bboxPartCom;             /* [GD_BBOX_BLOCKS][4]: FP64 m, m x, m y, m z */
    /* prefixSum scratch */
    uint32_t* scanTiles;             /* tile sums */
    /* per-node build records, maxNodes entries: first body, body count
       (octant in the parent in bits 29-31), parent, first child (-1 for a
       leaf) */
    int4* meta;
    /* compact walk arrays, maxNodes entries each */
    float4* wcom;                    /* centre of mass + mass */
    int2* wlink;                     /* .x first child (>= 0) or -(first body + 1) for a leaf,
                                        .y = (children or bodies) << 5 | level */
    float* wrad;                     /* mass radius: bound on the distance from the centre of
                                        mass to the node's bodies */
    int maxNodes;
    int noRecords;                   /* 1: builds leave the 80-byte node array alone */
    int lastBodies;                  /* bodies of the last octree build */
    /* moments about the COM, one 64-byte record per node (gd_mom_* below):
       second moments when quad != 0, third moments too when quad == 2 */
    float4* wmom;
    int quad;
    int momNodes;
} gd_builder;

static inline gd_sorter* gd_sorter_of(radixSorter* s) { return (gd_sorter*)s; }
static inline gd_builder* gd_builder_of(treeBuilder* b) { return (gd_builder*)b; }

/* The mass radii the production walks open cells by (b->wrad under the
 * mass-radius rule of gd_walk_set_opening_rule), or NULL for the size rule. */
const float* gd_walk_radii(const gd_builder* b);

/* ---- launchers shared between translation units (all on gd_stream()) */
void gd_launch_bbox(const float4* pos, const float3* lo, int n, gd_builder* b, float3* outMin,
                    float3* outMax);
void gd_launch_morton_dev(const float4* pos, int n, const gd_tree_info* info, uint64_t* keys,
                          uint32_t* idx);
void gd_launch_reorder(const float4* posIn, const float3* velIn, float4* posOut, float3* velOut,
                       const uint32_t* idx, int n);
void gd_launch_reorder_lo(const float4* posIn, const float3* velIn, const float3* loIn,
                          float4* posOut, float3* velOut, float3* loOut, const uint32_t* idx,
                          int n);
void gd_sort_pairs(gd_sorter* s, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                   uint32_t* vout, int n);
void gd_build_octree(gd_builder* b, const uint64_t* sortedKeys, const float4* sortedPos, int n,
                     node* nodes, int maxNodes, int records);
/* The 80-byte node records of the last gd_build_octree, from its per-node
 * meta and walk arrays (bit-identical to a build that writes them). */
void gd_build_node_records(gd_builder* b, const uint64_t* sortedKeys, node* nodes, int maxNodes);
void gd_exclusive_scan_u32(gd_builder* b, const uint32_t* in, uint32_t* out, int n);
/* When not null, gd_build_octree records this event between the topology
 * kernels and the centre-of-mass levels (build-stage timing). */
extern hipEvent_t gd_build_mid_event;

/* ---- wave helpers ------------------------------------------------------ */
#ifdef __HIP_DEVICE_COMPILE__
#define GD_DEVICE_ONLY 1
#endif

__device__ __forceinline__ unsigned gd_lane() { return __lane_id(); }
__device__ __forceinline__ unsigned long long gd_lanemask_lt() {
    unsigned l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

/* Moment record of a node: 64 bytes, one aligned 64-byte memory sector
 * (so a build writes whole sectors), 4 float4:
 *   (Sxx Syy Szz Sxy) (Sxz Syz Oxxx Oyyy) (Ozzz Oxxy Oxxz Oxyy) (Oyyz Oxzz Oyzz Oxyz)
 * S = sum m d d^T and O = sum m d d d about the centre of mass.  Readers get
 * the second moments as (Sxx Syy Szz Sxy), (Sxz Syz trace 0), the trace
 * summed as Sxx + Syy + Szz, and the third moments as (xxx yyy zzz xxy),
 * (xxz xyy yyz xzz), (yzz xyz 0 0). */
__device__ __forceinline__ void gd_mom_quad(const float4* __restrict__ wmom, int i, float4& q0,
                                            float4& q1) {
    const float4 a = wmom[4 * i], b = wmom[4 * i + 1];
    q0 = a;
    q1 = make_float4(b.x, b.y, a.x + a.y + a.z, 0.f);
}
__device__ __forceinline__ void gd_mom_oct(const float4* __restrict__ wmom, int i, float4& o0,
                                           float4& o1, float4& o2) {
    const float4 b = wmom[4 * i + 1], c = wmom[4 * i + 2], d = wmom[4 * i + 3];
    o0 = make_float4(b.z, b.w, c.x, c.y);
    o1 = make_float4(c.z, c.w, d.x, d.y);
    o2 = make_float4(d.z, d.w, 0.f, 0.f);
}
__device__ __forceinline__ void gd_mom_store(float4* __restrict__ wmom, int i, float4 q0,
                                             float sxz, float syz, float4 o0, float4 o1,
                                             float4 o2) {
    wmom[4 * i] = q0;
    wmom[4 * i + 1] = make_float4(sxz, syz, o0.x, o0.y);
    wmom[4 * i + 2] = make_float4(o0.z, o0.w, o1.x, o1.y);
    wmom[4 * i + 3] = make_float4(o1.z, o1.w, o2.x, o2.y);
}

/* The dispatcher hands workgroups to the 8 XCDs round-robin (block b runs on
 * XCD b % 8, each XCD with its own L2).  This bijection on [0, nblocks) deals
 * logical blocks to the XCDs in chunks of GD_XCD_CHUNK consecutive ones
 * (chunk c to XCD c % 8), so neighbouring work (groups adjacent in Morton
 * order, which read mostly the same tree nodes) shares an L2 while every XCD
 * still sees work from everywhere (whole contiguous ranges per XCD leave the
 * XCDs with the dense regions last to finish).  Blocks past the last whole
 * round of chunks keep their index. */
#define GD_XCDS 8
#define GD_XCD_CHUNK 16
__device__ __forceinline__ unsigned gd_xcd_block() {
    const unsigned b = blockIdx.x;
    const unsigned full = gridDim.x / (GD_XCDS * GD_XCD_CHUNK) * (GD_XCDS * GD_XCD_CHUNK);
    if (b >= full) return b;
    const unsigned x = b % GD_XCDS, i = b / GD_XCDS;
    return ((i / GD_XCD_CHUNK) * GD_XCDS + x) * GD_XCD_CHUNK + i % GD_XCD_CHUNK;
}

#endif
