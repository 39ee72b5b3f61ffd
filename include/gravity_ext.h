/*
 * gravity_ext.h -- entry points of libgravity.so with no counterpart in the
 * reference.  The reference-API functions (gravity_tree.h,
 * gravity_sorting.h, gravity_host.h) keep their signatures; everything the
 * engine adds is here, with a gd_ prefix.
 */
#ifndef GRAVITY_EXT_H
#define GRAVITY_EXT_H

#include "gravity_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error control ------------------------------------------------------
 * Default: an error prints file:line and a message, then exit(-1), as the
 * reference's CUDA_CHECK does (macros.h:8-17).  With exitOnError = 0 the
 * first error is kept (gd_last_error) and the failing call returns early;
 * outputs of that call are then unspecified. */
void gd_set_exit_on_error(int exitOnError);
const char* gd_last_error(void);   /* "" when none */
void gd_clear_error(void);

/* ---- device memory for callers without a HIP toolchain -------------------
 * Thin wrappers of hipMalloc / hipMemcpy / hipMemset on the current device;
 * copies are synchronous with respect to the null stream. */
void* gd_malloc(size_t bytes);
void gd_free(void* p);
void gd_memcpy_h2d(void* dst, const void* src, size_t bytes);
void gd_memcpy_d2h(void* dst, const void* src, size_t bytes);
void gd_memcpy_d2d(void* dst, const void* src, size_t bytes);
void gd_memset(void* p, int value, size_t bytes);
void gd_synchronize(void);
int gd_device_count(void);
/* Name of the current device ("gfx950 ..."); written into buf (NUL-terminated). */
void gd_device_name(char* buf, int bufLen);

/* ---- tree summary ---------------------------------------------------------
 * After buildTree / gd_build_tree_async: total node count, overflow flag
 * and the number of nodes whose mass summed to zero, read from the device
 * (synchronises). */
void gd_tree_summary(const treeState* tree, int* totalNodes, int* overflow, int* zeroMassNodes);

/*
 * buildTree without any host synchronisation: every kernel of the build is
 * queued on the null stream, the level tables stay on the device and the
 * host fields of *tree (levelOffsets, levelCount, deepestLevel) are NOT
 * updated (gd_tree_levels fetches them).  posLo, when not NULL, is the low
 * part of hi/lo positions (float3 per body) and is permuted with pos/vel
 * through the spare buffer spareLo, whose pointer is swapped with *posLo.
 */
void gd_build_tree_async(treeState* tree, float4** pos, float3** vel, float3** posLo,
                         float3** spareLo);

/* Copies the device level tables into tree->levelOffsets / levelCount /
 * deepestLevel (synchronises). */
void gd_tree_levels(treeState* tree);

/* ---- production walk ------------------------------------------------------ */
typedef struct {
    unsigned long long interactions;   /* accepted body-node interactions */
    unsigned long long opened;         /* node tests that opened the node */
    unsigned long long overflowGroups; /* groups re-walked by the fallback */
} gd_walk_stats;

/*
 * Wavefront walk: one 64-body group per wavefront, the group's bounding box
 * tested against 64 nodes at a time (size^2 < theta^2 d^2, d the distance
 * from the box to the node's centre of mass, size = bbox cube extent / 2^level),
 * accepted nodes gathered in an LDS interaction list and evaluated by all 64
 * bodies; per-body sums are FP32, compensated (Kahan) across list flushes.
 * Needs a tree built by buildTree or gd_build_tree_async.  stats may be
 * NULL; when given, counters are accumulated on the device and copied out
 * (synchronises).  phi (may be NULL) receives the softened potential.
 */
void gd_walk(const treeState* tree, const float4* pos, float3* acc, float* phi, float eps2,
             float G, float theta, int starCount, gd_walk_stats* stats);

/* Time the wavefront-walk kernel of every later gd_walk call with HIP events
 * on the null stream (enable = 1) or stop (0).  gd_walk_kernel_ms returns the
 * last timed kernel's duration in ms (synchronises), or -1 if none. */
void gd_walk_timing(int enable);
float gd_walk_kernel_ms(void);

/* ---- hi/lo positions -------------------------------------------------------
 * Kick-drift with positions held as hi (float4 .xyz) + lo (float3):
 * v += a dt; (hi, lo) += v dt with an error-free two-sum, renormalised so
 * that hi = fl(hi + lo). */
void gd_integrate_hilo(float4* pos, float3* posLo, float3* vel, const float3* acc,
                       int starCount, float dt);

/* ---- FP64 direct-sum judge ------------------------------------------------
 * acc[3k..3k+2] (device, double) = sum_{j != sample[k]} -G m_j (x_i - x_j) /
 * (r^2 + eps2)^{3/2} in FP64 over all starCount bodies.  posLo may be NULL.
 * sample: device int array of ns body indices. */
void gd_direct_sum_fp64(const float4* pos, const float3* posLo, int starCount,
                        const int* sample, int ns, double eps2, double G, double* acc);

/* ---- FP64 diagnostics -----------------------------------------------------
 * out[0..6] (host) = kinetic energy, momentum x/y/z, mass, and the
 * potential energy 0.5 sum m_i phi_i when phi != NULL (else 0), all reduced in
 * FP64 on the device; out[6] = total energy. */
void gd_diagnostics(const float4* pos, const float3* vel, const float* phi, int starCount,
                    double* out);

/* ---- timing ----------------------------------------------------------------
 * HIP events on the null stream, the stream every entry point launches on. */
void* gd_event_create(void);
void gd_event_record(void* ev);
float gd_event_elapsed_ms(void* start, void* stop);   /* synchronises on stop */
void gd_event_destroy(void* ev);

#ifdef __cplusplus
}
#endif

#endif
