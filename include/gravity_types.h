/*
 * gravity_types.h -- body, node and tree-state layouts of the Barnes-Hut engine.
 *
 * Every struct here keeps the byte layout of the reference's types.h so that a
 * C main loop written against the reference links against this library
 * unchanged:
 *   sVec4, sVec3, RGB ..................... types.h:7-17
 *   radixSorter ........................... types.h:19-23   (24 B)
 *   worldBox .............................. types.h:25-34   (20 B)
 *   node .................................. types.h:52-62   (80 B)
 *   treeBuilder ........................... types.h:64-73   (56 B)
 *   treeState ............................. types.h:75-90   (304 B)
 *   galaxyDiskParameters .................. types.h:92-100
 *   sorterPair ............................ types.h:102-105
 *
 * float3 / float4 are HIP's vector types when compiled by hipcc; plain C and
 * C++ callers get structs with the same size and alignment (12 B and 16 B,
 * 16-byte aligned), which is what the reference got from cuda_runtime.h.
 *
 * No HIP header is needed by a caller: every pointer in these structs is an
 * opaque device address.
 */
#ifndef GRAVITY_TYPES_H
#define GRAVITY_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#else
typedef struct { float x, y, z; } float3;
typedef struct __attribute__((aligned(16))) { float x, y, z, w; } float4;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Position + mass of one body (types.h:7-9). */
typedef struct __attribute__((aligned(16))) {
    float x, y, z, m;
} sVec4;

/* Velocity / acceleration of one body (types.h:11-13). */
typedef struct {
    float x, y, z;
} sVec3;

/* Per-body colour written in the trajectory header (types.h:15-17). */
typedef struct {
    uint8_t R, G, B;
} RGB;

/*
 * Sorter handle (types.h:19-23).  tempStorage is the device workspace of the
 * radix sort (digit histograms, look-back tile status words, tile counters);
 * tempStorageBytes its size; capacity the largest count one sort accepts.
 */
typedef struct {
    void* tempStorage;
    size_t tempStorageBytes;
    int capacity;
} radixSorter;

/*
 * Integer-grid mapping of the bounding cube (types.h:25-34):
 * q = (uint32)((x - min) * scale), scale = (2^21 - 1) / extent, extent =
 * max(width, height, depth).
 */
typedef struct {
    float minX, minY, minZ;
    float extent;
    float scale;
} worldBox;

/*
 * One octree node (types.h:52-62), 80 bytes:
 *   @0  nodePathFromRoot  Morton key >> (63 - 3*treeLevel)
 *   @8  parentIndex       index of the parent in the node array (root: -1 here)
 *   @12 firstParticleIndex first body (sorted order) inside the node
 *   @16 particleCount     number of bodies inside the node
 *   @20 child[8]          child node index per octant, -1 when empty
 *   @52 treeLevel         0 = root ... 21
 *   @56 (8 bytes padding; the reference never writes them)
 *   @64 massData          centre of mass (x, y, z) and total mass (w)
 * The array is level-major: the root at 0, then every level-1 node, then
 * level 2, ...; inside a level, nodes ascend by firstParticleIndex.
 */
typedef struct {
    uint64_t nodePathFromRoot;
    int parentIndex;
    int firstParticleIndex;
    int particleCount;
    int child[8];
    int treeLevel;
    float4 massData;
} node;

/*
 * Octree builder handle (types.h:64-73).  The four flag/offset arrays hold N
 * u32 each and are the per-level node-start flags and their exclusive scan
 * (the ping-pong pair of the level-wise build); tempStorage is the device
 * workspace of the scan.
 */
typedef struct {
    void* tempStorage;
    size_t tempStorageBytes;
    int capacity;

    uint32_t* flagsPing;
    uint32_t* flagsPong;
    uint32_t* offsetsPing;
    uint32_t* offsetsPong;
} treeBuilder;

/*
 * Everything one octree needs (types.h:75-90).  Device pointers: nodes,
 * mortonIn/Out, indexIn/Out, boundingData[0..3] (per-block min, per-block
 * max, global min, global max corners), sortedPosMassVals, sortedVelVals.
 * Host fields: levelOffsets / levelCount (node-array start and node count of
 * each level), deepestLevel, treeBlocks (256-thread blocks covering N),
 * starCount.  The struct is passed by value to calculateAcceleration.
 */
typedef struct {
    node* nodes;
    int maxNodes;
    uint64_t* mortonIn; uint64_t* mortonOut;
    uint32_t* indexIn; uint32_t* indexOut;
    float3* boundingData[4];
    radixSorter* sorter;
    treeBuilder* builder;
    int levelOffsets[22];
    int levelCount[22];
    int deepestLevel;
    int treeBlocks;
    int starCount;
    float4* sortedPosMassVals;
    float3* sortedVelVals;
} treeState;

/* Initial-condition parameters of one disk galaxy (types.h:92-100). */
typedef struct {
    float scaleLength;
    float scaleThickness;

    float totalDiskMass;
    sVec3 rotationAngles;          /* yaw, pitch, roll */
    sVec3 galacticCentreLocation;
    sVec3 galacticNetVelocity;
} galaxyDiskParameters;

/* (position, original index) pair used by the ordered frame writer (types.h:102-105). */
typedef struct {
    float4* positionVal;
    uint32_t* originalIndex;
} sorterPair;

#ifdef __cplusplus
}  /* extern "C" */
#define GRAVITY_STATIC_ASSERT static_assert
#else
#define GRAVITY_STATIC_ASSERT _Static_assert
#endif

GRAVITY_STATIC_ASSERT(sizeof(float3) == 12, "float3 must be 12 bytes");
GRAVITY_STATIC_ASSERT(sizeof(float4) == 16, "float4 must be 16 bytes");
GRAVITY_STATIC_ASSERT(sizeof(sVec4) == 16, "sVec4 layout");
GRAVITY_STATIC_ASSERT(sizeof(radixSorter) == 24, "radixSorter layout");
GRAVITY_STATIC_ASSERT(sizeof(worldBox) == 20, "worldBox layout");
GRAVITY_STATIC_ASSERT(sizeof(node) == 80, "node layout");
GRAVITY_STATIC_ASSERT(offsetof(node, parentIndex) == 8, "node layout");
GRAVITY_STATIC_ASSERT(offsetof(node, child) == 20, "node layout");
GRAVITY_STATIC_ASSERT(offsetof(node, treeLevel) == 52, "node layout");
GRAVITY_STATIC_ASSERT(offsetof(node, massData) == 64, "node layout");
GRAVITY_STATIC_ASSERT(sizeof(treeBuilder) == 56, "treeBuilder layout");
GRAVITY_STATIC_ASSERT(sizeof(treeState) == 304, "treeState layout");
GRAVITY_STATIC_ASSERT(offsetof(treeState, levelOffsets) == 96, "treeState layout");
GRAVITY_STATIC_ASSERT(offsetof(treeState, sortedPosMassVals) == 288, "treeState layout");

#endif /* GRAVITY_TYPES_H */
