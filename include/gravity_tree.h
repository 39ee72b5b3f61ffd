/*
 * gravity_tree.h -- the reference's tree.h API (tree.h:16-46), C linkage.
 *
 * The reference declares its ten kernels here and launches them with
 * <<<grid, block>>>.  A C caller cannot, so each kernel is exported as a host
 * launcher with the kernel's own argument list; the launcher picks the
 * reference's launch geometry (noted per function) and runs on the null
 * stream.  All array arguments are device pointers.
 */
#ifndef GRAVITY_TREE_H
#define GRAVITY_TREE_H

#include "gravity_types.h"

#define BOXTHREADS 256      /* tree.h:16 */
#define BOXBLOCKS 256       /* tree.h:17 */
#define WARPTHREADS 32      /* tree.h:18: size of a walk group */

#ifdef __cplusplus
extern "C" {
#endif

/* tree.cu:347-367: keys[i] = 63-bit Morton key of body i quantised by rootBox,
 * originalIndex[i] = i.  ceil(NStars/256) x 256. */
void mortonEncode(const float4* posMassVals, uint64_t* mortonCodes, uint32_t* originalIndex,
                  int NStars, worldBox rootBox);

/* tree.cu:414-456: BOXBLOCKS per-block min/max partials of N positions. */
void localMinMaxFinder(const float4* posMassVals, unsigned int N, float3* minCorner,
                       float3* maxCorner);

/* tree.cu:459-490: reduce count (<= BOXTHREADS) partials into outMin/outMax. */
void globalMinMaxReducer(const float3* minCorner, const float3* maxCorner, int count,
                         float3* outMin, float3* outMax);

/* tree.cu:50-78: nodeFlags[i] = 1 iff body i shares its level-(nodeLevel-1)
 * cell with a neighbour and starts a new level-nodeLevel cell. */
void identifyNodesAtLevel(uint64_t* mortonCodes, int nodeLevel, int starCount, uint32_t* nodeFlags);

/* tree.cu:205-235: writes the level-`level` nodes and links them to their
 * parents (`flags`/`offsets` of this level, `previousFlags`/`offsetsPrev` of
 * the level above). */
void createNodes(const uint64_t* mortonCodes, const uint32_t* flags, const uint32_t* previousFlags,
                 const uint32_t* offsets, const uint32_t* offsetsPrev, node* nodes, int level,
                 int arrayLevelOffset, int prevArrayLevelOffset, int starCount, int maxNodes);

/* tree.cu:181-203: particleCount of the levelNodeCount nodes starting at
 * levelStartIndex, from the next sibling or the parent's end. */
void updateTreeParticles(node* nodes, int levelStartIndex, int levelNodeCount, int starCount);

/* tree.cu:89-179: centre of mass of the levelNodes nodes of one level
 * (children must be final).  Nodes of zero total mass get zeros. */
void updateTreeMass(node* nodes, const float4* positionMassVals, int level, int levelNodes,
                    int levelStartIndex);

/*
 * tree.cu:562-676: tree walk, one 32-body group (WARPTHREADS) per vote:
 * the group descends a node when any of its bodies fails
 * size^2 < theta^2 r^2 on a non-leaf, size = (bbox x extent) / 2^level;
 * otherwise every body adds -G m d / (r^2 + eps^2)^{3/2}.  accelVals[i] is
 * written for i < starCount.  Two groups share one 64-wide wavefront, each
 * with its own stack and vote.
 */
void calculateAcceleration(const float4* positionMassArray, treeState tree, float3* accelVals,
                           float antiSingularitySquared, float GVal, float theta, int starCount);

/* tree.cu:678-703: v += a dt, then x += v dt, for starCount bodies. */
void positionUpdater(float4* positionMassArray, float3* velocityVals, float3* accelerationVals,
                     int starCount, float timeStep);

/* tree.cu:241-284: builder for the per-level exclusive scan, with the four
 * flag/offset arrays of maxCount u32 each.  maxCount <= 0 is an error. */
treeBuilder* sumCreate(int maxCount);

/* tree.cu:286-302 (declared tree.h:38 with C++ linkage; exported here with C
 * linkage): offsets = exclusive scan of flags[0..maxCount). */
void prefixSum(treeBuilder* builder, uint32_t* flags, uint32_t* offsets, int maxCount);

/* tree.cu:304-318 (declared in no reference header). */
void sumDestroyer(treeBuilder* sum);

#ifdef __cplusplus
}
#endif

#endif
