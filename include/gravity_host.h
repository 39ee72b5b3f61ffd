/*
 * gravity_host.h -- the reference's host orchestration API, C linkage:
 * the functions of basicGravity.cu that its main loop calls (declared in no
 * reference header), main.h's launch constants and units.h's globals.
 *
 * A C main loop written against the reference includes this header (with
 * gravity_tree.h and gravity_sorting.h) and links libgravity.so.
 */
#ifndef GRAVITY_HOST_H
#define GRAVITY_HOST_H

#include "gravity_types.h"
#include "gravity_tree.h"
#include "gravity_sorting.h"

#define THREADCOARSENING 2      /* main.h:25 (unused by the reference) */
#define THREADPERBLOCK 256      /* main.h:26 */

#ifdef __cplusplus
extern "C" {
#endif

/* units.c:3-4: G = 1 (mass 1e10 Msun, length 1 kpc), default opening angle. */
extern float G;
extern float theta;

/*
 * basicGravity.cu:5-77.  Allocates the sorter, the builder, maxNodes = 8 N
 * nodes, both key and index arrays, the sorted position / velocity buffers
 * (paddedStarCount, zeroed) and the bbox scratch; indexIn = iota.
 */
void initialiseTree(treeState* state, int starCount, int paddedStarCount, int integrateBlocks);

/*
 * basicGravity.cu:79-113.  Global bbox of starCount positions into
 * outputMin/outputMax (device, one float3 each) via BOXBLOCKS partials in
 * minCorner/maxCorner, then the quantisation box on the host.
 */
void createWorldBox(const float4* cudaPositionMassVals, const int starCount,
                    float3* cudaMinCorner, float3* cudaMaxCorner,
                    float3* cudaOutputMin, float3* cudaOutputMax, worldBox* outputWorldBox);

/*
 * basicGravity.cu:116-325.  Builds the octree of *pos: bbox, keys, stable
 * sort, reorder of pos/vel into tree's sorted buffers and a swap of the
 * caller's pointers with them (the caller must use the returned pointers),
 * then topology and centre of mass.  On return nodes, mortonOut, indexOut,
 * levelOffsets, levelCount and deepestLevel describe the new tree (the
 * node array equals the reference's level-wise build; the root's
 * parentIndex is -1).  An overflowing node array is an error.
 */
void buildTree(treeState* tree, float4** cudaPosMasVals, float3** cudaVelVals);

/* basicGravity.cu:351-366 */
int roundNumberToMultiple(int number, int multiple);

/* basicGravity.cu:374-427: walk groups of WARPTHREADS, 256-thread blocks for
 * the build and the integrator, N padded to a multiple of WARPTHREADS. */
void blockNumCalculator(int* accelerationBlocks, int* treeBuilderBlocks, int* integrateBlocks,
                        int* paddedStarCount, int* integrationThreadsPerBlock,
                        int* treeBuilderThreadsPerBlock, int starCount);

#ifdef __cplusplus
}
#endif

#endif
