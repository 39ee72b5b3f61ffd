/*
 * gravity_sorting.h -- the reference's sorting.h API (sorting.h:12-23),
 * C linkage, implemented by hand-written gfx950 kernels in libgravity.so.
 */
#ifndef GRAVITY_SORTING_H
#define GRAVITY_SORTING_H

#include "gravity_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Replaces sorting.cu:15-63.  Allocates a sorter for up to maxCount pairs
 * (device workspace: ping-pong key/value buffers, digit histograms, look-back
 * status words).  maxCount <= 0 is an error.
 */
radixSorter* sorterCreate(int maxCount);

/*
 * Replaces sorting.cu:65-79 (CUB SortPairs over bits [0,64)).  Stable
 * ascending LSD radix sort of numCount (u64 key, u32 value) pairs from
 * keysIn/valsIn (left untouched) into keysOut/valsOut.  All pointers are
 * device pointers; keysOut/valsOut must not alias the inputs.  A NULL sorter
 * or numCount outside [1, capacity] is an error.  Runs on the library's
 * stream, which synchronises with the null stream.
 */
void radixSortPairs(radixSorter* sorter, const uint64_t* keysIn, uint64_t* keysOut,
                    const uint32_t* valsIn, uint32_t* valsOut, int numCount);

/* Replaces sorting.cu:81-92.  NULL is an error. */
void sorterDestroyer(radixSorter* s);

/*
 * Launcher for the reference kernel reorderParticles (sorting.cu:99-107,
 * declared at sorting.h:12): posOut[i] = posIn[idx[i]], velOut[i] =
 * velIn[idx[i]] for i < starCount.  Device pointers, the library's stream.
 */
void reorderParticles(const float4* posMassValsIn, const float3* velValsIn,
                      float4* posMassValsOut, float3* velValsOut,
                      uint32_t* originalIndex, int starCount);

#ifdef __cplusplus
}
#endif

#endif
