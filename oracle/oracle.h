/*
 * oracle.h -- serial C restatement of the reference's Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine
 * is compared against (and the timed CPU baseline in bench.py).  Nothing in
 * the product path may link, load or call it.
 *
 * Each function restates one reference kernel or host routine, one body /
 * thread at a time, in the reference's order.  The reference arithmetic is
 * defined here as the documented operation order below (the reference binary
 * cannot be built, so its compiler's contraction choices are not observable):
 *   - no implicit contraction (built with -ffp-contract=off);
 *   - fmaf exactly where a product feeds a running sum:
 *       centre of mass  s = fmaf(x, m, s)                 (tree.cu:114-116,161-163)
 *       walk distance   d2 = fmaf(dz,dz, fmaf(dy,dy, dx*dx))  (tree.cu:611)
 *       walk accumulate a = fmaf(mag, d/dist, a)          (tree.cu:666-668)
 *       kick, drift     v = fmaf(a, dt, v); x = fmaf(v, dt, x) (tree.cu:695-702)
 *   - IEEE division and square root; node size = extent / 2^level computed
 *     in double and rounded to float (tree.cu:527).
 *
 * Which check pins each function is stated at the function.
 */
#ifndef GRAVITY_ORACLE_H
#define GRAVITY_ORACLE_H

#include "gravity_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Exact per-axis min/max of positions[0..n) (tree.cu:414-490). */
void or_bbox(const float4* pos, int n, float3* outMin, float3* outMax);

/* worldBox from the two corners (basicGravity.cu:90-112). */
void or_world_box(const float3* mn, const float3* mx, worldBox* box);

/* 63-bit Morton key and iota payload (tree.cu:332-367). */
void or_morton(const float4* pos, int n, const worldBox* box, uint64_t* keys, uint32_t* idx);

/* Stable ascending sort of (key, value) pairs over all 64 key bits
 * (sorting.cu:65-79; CUB SortPairs semantics). */
void or_sort_pairs(const uint64_t* keysIn, uint64_t* keysOut,
                   const uint32_t* valsIn, uint32_t* valsOut, int n);

/* Gather positions and velocities into sorted order (sorting.cu:99-107). */
void or_reorder(const float4* posIn, const float3* velIn, float4* posOut, float3* velOut,
                const uint32_t* idx, int n);

/* Level-wise octree topology from sorted keys (basicGravity.cu:158-169,229-310;
 * tree.cu:50-78,181-235).  Fills nodes[0..total), levelOffsets / levelCount for
 * levels 0..deepest (+ the empty terminating level when deepest < 21) and
 * *deepest.  Returns 0, or -1 when the node array would overflow maxNodes
 * (basicGravity.cu:282).  The root's parentIndex is set to -1. */
int or_build_topology(const uint64_t* sortedKeys, int n, node* nodes, int maxNodes,
                      int* levelOffsets, int* levelCount, int* deepest);

/* Bottom-up centre of mass, deepest level first (basicGravity.cu:317-319,
 * tree.cu:89-179).  Returns the number of nodes whose total mass was zero. */
int or_tree_mass(node* nodes, const float4* pos, const int* levelOffsets,
                 const int* levelCount, int deepest);

/* Warp-voted tree walk over 32-body groups (tree.cu:501-676).  bmin/bmax are
 * the global bbox corners (only their x extent sizes nodes, tree.cu:506).
 * Groups g in [g0, g1) with g % gstride == 0 are walked; acc is written only
 * for their bodies.  Returns the number of accepted body-node interactions
 * of the valid lanes; *opened receives the number of valid-lane node tests
 * that voted to descend. */
long long or_walk(const float4* pos, int n, const node* nodes, const float3* bmin,
                  const float3* bmax, float eps2, float G, float theta, float3* acc,
                  int g0, int g1, int gstride, long long* opened);

/* or_walk over every group, also recording the nodes each valid body
 * accepted: bodyCount[i] (may exceed perBody) and lists[i * perBody + k],
 * in traversal order.  variant 0 is or_walk's rule.  Variants 1-5 are
 * deliberate mistakes for the pinning tests, which must reject each:
 *   1 per-lane acceptance (no 32-body vote)   2 node size from the widest axis
 *   3 '<=' in the opening test                4 invalid lanes vote
 *   5 leaf rule without its no-children clause
 * Returns or_walk's count, or -1 for an unknown variant. */
long long or_walk_traced(const float4* pos, int n, const node* nodes, const float3* bmin,
                         const float3* bmax, float eps2, float G, float theta, float3* acc,
                         long long* opened, int variant, int* bodyCount, int* lists,
                         int perBody);

/* Semi-implicit Euler kick-drift (tree.cu:678-703). */
void or_integrate(float4* pos, float3* vel, const float3* acc, int n, float dt);

/* FP64 softened direct sum on the bodies listed in sample[0..ns):
 * acc[3k..3k+2] = sum_{j != sample[k]} -G m_j (x_i - x_j) / (r^2 + eps2)^{3/2}. */
void or_direct_sum_fp64(const float4* pos, int n, const int* sample, int ns,
                        double eps2, double G, double* acc);

/*
 * One full reference step on host arrays (basicGravity.cu:774-783):
 * buildTree (bbox, keys, sort, reorder with pointer swap, topology, mass),
 * walk, kick-drift.  pos/vel/spare arrays are swapped exactly like the
 * reference swaps the caller's pointers with treeState's sorted buffers, so
 * *pos / *vel point at the sorted arrays on return.  keys/idx are scratch of
 * 2n each, nodes holds maxNodes.  Returns 0 or -1 on node overflow.
 */
int or_step(float4** pos, float3** vel, float4** posSpare, float3** velSpare, float3* acc,
            int n, uint64_t* keysIn, uint64_t* keysOut, uint32_t* idxIn, uint32_t* idxOut,
            node* nodes, int maxNodes, int* levelOffsets, int* levelCount, int* deepest,
            float eps2, float G, float theta, float dt);

#ifdef __cplusplus
}
#endif

#endif
