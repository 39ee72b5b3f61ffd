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

/* The root-level centre of mass in FP64: out[0..2] = sum m x / sum m over
 * all bodies (hi + lo positions when the build had them), out[3] = sum m,
 * summed in FP64 in a fixed order by the build's bounding-box pass (the
 * node array keeps the reference's FP32 centre of mass).  Synchronises. */
void gd_tree_root_fp64(const treeState* tree, double* out);

/*
 * buildTree without any host synchronisation: every kernel of the build is
 * queued on the library's stream, the level tables stay on the device and the
 * host fields of *tree (levelOffsets, levelCount, deepestLevel) are NOT
 * updated (gd_tree_levels fetches them).  posLo, when not NULL, is the low
 * part of hi/lo positions (float3 per body) and is permuted with pos/vel
 * through the spare buffer spareLo, whose pointer is swapped with *posLo.
 */
void gd_build_tree_async(treeState* tree, float4** pos, float3** vel, float3** posLo,
                         float3** spareLo);

/* Whether later gd_build_tree_async builds also write the reference's 80-byte
 * node array tree->nodes (enable = 1, the default) or only the compact walk
 * arrays the engine's walks read (0; tree->nodes then keeps what it held).
 * buildTree always writes the node array.  gd_write_node_records writes it
 * for the last build, on the library's stream, byte for byte as a build
 * with records on would have (it reads the build's keys, tree->mortonOut,
 * so call it before anything else writes them). */
void gd_set_node_records(treeState* tree, int enable);
void gd_write_node_records(treeState* tree);

/* Time every later gd_build_tree_async / buildTree with HIP events on the
 * library's stream (enable = 1) or stop (0).  gd_build_phase_ms fills ms[0..5]
 * with the last build's bbox, keys, sort, reorder, topology (node counts,
 * level scan, node records) and centre-of-mass (all levels, with the
 * moments) times and returns 0, or -1 when nothing was timed (synchronises). */
void gd_build_timing(int enable);
int gd_build_phase_ms(float* ms);

/* Makes every later build also compute each node's moments about its centre
 * of mass: enable = 1 the second moments S = sum m (x-c)(x-c)^T (used by the
 * quadrupole walk), 2 also the third moments O = sum m (x-c)(x-c)(x-c) (the
 * learned walk's inputs); 0 stops both.  Both live in one 64-byte record per
 * node of capacity (allocated on the first call that enables them). */
void gd_set_quadrupole(treeState* tree, int enable);

/* Copies the third moments of nodes [0, count) to host memory, 3 float4 per
 * node: Oxxx Oyyy Ozzz Oxxy, Oxxz Oxyy Oyyz Oxzz, Oyzz Oxyz 0 0
 * (gd_set_quadrupole(tree, 2) before the build).  Synchronises. */
void gd_tree_third_moments(const treeState* tree, void* oct, int count);

/* Copies the compact walk records of nodes [0, count) to host memory:
 * com (float4: centre of mass, mass), link (int2: first child or -(first
 * body + 1); count << 5 | level) and, when not NULL and moments are on, quad
 * (2 float4: Sxx Syy Szz Sxy, Sxz Syz trace 0).  Synchronises. */
void gd_tree_walk_arrays(const treeState* tree, void* com, void* link, void* quad, int count);

/* Copies the mass radii of nodes [0, count) to host memory: for a leaf the
 * largest distance from its centre of mass to its bodies (0 for one body),
 * for an internal node the largest |c_child - c| + radius_child over its
 * children (float32, sqrt((dx dx + dy dy) + dz dz) + radius), so a bound on
 * the distance from c to every body below.  Every build writes them.
 * Synchronises. */
void gd_tree_mass_radii(const treeState* tree, float* rad, int count);

/* Copies the device level tables into tree->levelOffsets / levelCount /
 * deepestLevel (synchronises). */
void gd_tree_levels(treeState* tree);

/* ---- production walk ------------------------------------------------------ */
typedef struct {
    unsigned long long interactions;     /* accepted body-node interactions */
    unsigned long long opened;           /* node tests that opened the node */
    unsigned long long overflowGroups;   /* groups whose first-pass list did not fit */
    unsigned long long cellInteractions; /* of the interactions, those with internal
                                            cells evaluated as quadrupoles (+ network) */
    unsigned long long bodyGroups;       /* of those, groups walked body by body */
    unsigned long long networkEvals;     /* learned walk: (group or body, cell) pairs
                                            evaluated by the network */
    unsigned long long farCells;         /* (group, cell) pairs summed through the
                                            group's local expansion (far field) */
    unsigned long long farInteractions;  /* of cellInteractions, the body-cell pairs
                                            those expansions stand for */
} gd_walk_stats;

/*
 * Wavefront walk: one 64-body group per wavefront, the group's bounding box
 * tested against 64 nodes at a time (size^2 < theta^2 d^2, d the distance
 * from the box to the node's centre of mass, size = bbox cube extent / 2^level),
 * accepted nodes written to the group's interaction list in device memory
 * and evaluated by all 64 bodies; per-body sums are FP32, compensated
 * (Kahan) across 64-entry chunks.  A group whose list does not fit its
 * capacity is walked again with a larger list (second pass); one that still
 * does not fit is walked body by body (the same acceptance test against the
 * body's own position, no list).  With the second moments built
 * (gd_set_quadrupole), accepted internal cells are evaluated as quadrupoles.
 * Needs a tree built by buildTree or gd_build_tree_async.  stats may be
 * NULL; when given, counters are accumulated on the device and copied out
 * (synchronises).  phi (may be NULL) receives the softened potential.  When
 * the build ran out of nodes (the reference stops there, basicGravity.cu:282)
 * the walk reads no tree: it fails with "Node array overflow" and leaves acc
 * unwritten.
 */
void gd_walk(const treeState* tree, const float4* pos, float3* acc, float* phi, float eps2,
             float G, float theta, int starCount, gd_walk_stats* stats);

/* Interaction-list capacities (entries of 4 bytes).  entries: per 64-body
 * group in the first pass.  0 (the default) packs the first pass's lists
 * back to back, each group getting 1.5 times its previous walk's list length
 * and at least 4096 entries; groups found spread out (walked body by body)
 * skip the list passes, re-checked every 32 walks.  A value > 0 gives every
 * group exactly that many entries and no skipping.  Either call forgets the
 * recorded lengths.  secondPass: per
 * group in the second pass (0: 262144; at most 8 GiB of lists at a time).
 * Returns the previous first-pass setting. */
int gd_walk_set_list_capacity(int entries, int secondPass);

/* Asynchronous walks (enable = 1): gd_walk and gd_walk_learned then make
 * no host synchronisation unless statistics are asked for -- the second
 * pass runs on fixed grids for up to 256 groups whose count the kernels read
 * from device memory (groups beyond them take the per-body walk), the
 * per-body walk loops over a device-held count, and the walk's status comes
 * back to pinned host memory without waiting.  The next walk (or
 * gd_walk_poll) reads that status when it has arrived: it sizes the list
 * pool and reports a tree that did not fit its node array ("Node array
 * overflow"), one walk late.  Accelerations are those of the synchronous walk
 * whenever at most 256 groups overflow the first pass.  Returns the previous
 * setting; switching off waits for the last status. */
int gd_walk_set_async(int enable);
/* Waits for the last asynchronous walk's status and applies it; returns 0,
 * or -1 after reporting an error (a tree that did not fit). */
int gd_walk_poll(void);

/* Groups of the list walks.  Each run of 64 bodies in Morton order is split
 * in two at its largest key jump; each half has its own interaction list
 * (seen through its own, smaller bounding box) and the two lists are
 * evaluated side by side by one wavefront.  ratio < 0 (the default): every
 * run with two distinct keys is split; 0: whole runs; 0 < ratio <= 1: a run
 * is split only when the halves' boxes hold less than `ratio` of the run's
 * box volume.  The training samples (gd_corrector_samples) take the same
 * groups.  Returns the previous setting. */
float gd_walk_set_group_split(float ratio);

/* Far field of the list walks (quadrupole and learned walks without
 * potential).  With 0 < ratio < 1, a cell of a group's list whose centre of
 * mass is at distance R from the centre of the group's bounding box, with
 * h / R < ratio (h the box's half diagonal), is not evaluated body by body:
 * its monopole and quadrupole field is summed, once per group, into a local
 * (Taylor) expansion about the box centre, third order for the monopoles and
 * second order for the second moments, which each body then evaluates at
 * its own position (truncation O((h/R)^4) and O((h/R)^3) of those terms).
 * Leaves, and in the learned walk the cells the network corrects, stay body
 * by body.  The interaction list is unchanged.  0 (the default) evaluates
 * every cell body by body.  Second-pass and per-body walks always do.
 * Returns the previous setting. */
float gd_walk_set_far_field(float ratio);

/* Opening rule of the production walks (list walks, second pass, per-body
 * walk and the training samples; not the reference walk calculateAcceleration).
 * 0 (the default): an internal cell is accepted for a box when
 * size^2 < theta^2 d^2, size = extent / 2^level and d the distance from the
 * box to the cell's centre of mass.  1: b^2 < theta^2 d^2 with b the cell's
 * mass radius (gd_tree_mass_radii), a bound on the distance from its centre
 * of mass to its bodies, so that a cell whose bodies lie close to their
 * centre of mass is accepted nearer; the same theta then means a different
 * (smaller) error.  Under rule 1 the learned walk and its training samples
 * also take the mass radius as the cell measure s of the network's inputs,
 * of its correction test s / r > minRatio and of the target's scale (a
 * network trained under one rule is for that rule).  Returns the previous
 * rule, or -1 after reporting an error (a rule other than 0 or 1). */
int gd_walk_set_opening_rule(int rule);

/* Time every later gd_walk / gd_walk_learned call with HIP events on the
 * library's stream (enable = 1) or stop (0).  The walk is a traversal kernel that
 * writes each group's interaction list, an evaluation kernel (the dominant
 * one) and, for groups whose list did not fit, the per-body walk.
 * gd_walk_kernel_ms returns the last evaluation kernel's duration in ms;
 * gd_walk_phase_ms fills ms[0..2] with traversal, evaluation and per-body
 * walk times and returns 0 (or -1 when nothing was timed).  Both synchronise. */
void gd_walk_timing(int enable);
float gd_walk_kernel_ms(void);
int gd_walk_phase_ms(float* ms);

/* ---- learned far-field corrector ------------------------------------------
 * Network 8 -> 64 -> 64 -> 64 -> 3 with GELU, BF16 weights/activations and
 * FP32 accumulation on the matrix cores.  Parameters (FP32, row-major
 * [out][in]): W1[64][8] @0, b1 @512, W2[64][64] @576, b2 @4672,
 * W3[64][64] @4736, b3 @8832, W4[3][64] @8896, b4 @9088 (9091 values).
 * Inputs per (point, cell), with u the unit vector from the point to the
 * cell's COM, r the distance, s the cell's edge, m its mass, O its third
 * moments about the COM, w_i = O_ijk u_j u_k and T_i = O_ijj: u (3), s / r,
 * (7.5 w - 1.5 T) / (m s^3) (3), u.w / (m s^3).  Output: the correction of
 * the cell's acceleration beyond its monopole and quadrupole terms, in units
 * of G m s^3 / r^5 (the octupole term is a linear function of the inputs in
 * these units).  X (rows x 8), Y (rows x 3) and order are device arrays. */
#define GD_MLP_PARAMS 9091
typedef struct gd_mlp gd_mlp;
gd_mlp* gd_mlp_create(const float* hostParams);          /* NULL: all zero */
void gd_mlp_destroy(gd_mlp* net);
void gd_mlp_get_params(const gd_mlp* net, float* hostParams);
void gd_mlp_set_params(gd_mlp* net, const float* hostParams);
/* Y = net(X) for `rows` feature rows. */
void gd_mlp_forward(const gd_mlp* net, const float* X, float* Y, long long rows);
/* Gradient of mean_b |net(X[order[b]]) - Y[order[b]]|^2 over b < B into
 * hostGrad (may be NULL); returns the loss.  Synchronises. */
float gd_mlp_gradient(gd_mlp* net, const float* X, const float* Y, const int* order, long long B,
                      float* hostGrad);
/* The same gradient followed by one Adam step (beta 0.9 / 0.999, eps 1e-8,
 * bias-corrected) on the FP32 parameters; returns the loss. */
float gd_mlp_train_step(gd_mlp* net, const float* X, const float* Y, const int* order,
                        long long B, float lr);
void gd_mlp_adam_state(const gd_mlp* net, float* hostM, float* hostV);
/* Training samples from the current tree (built with second and third
 * moments, gd_set_quadrupole(tree, 2)): for each target body (device
 * indices, sorted order), every internal cell that the traversal of the
 * body's group (as the list walks form groups, gd_walk_set_group_split) at
 * opening angle theta accepts (under the walks' opening rule) and whose
 * s / distance from the group's centre exceeds minRatio (s as in
 * gd_walk_learned: the cells it corrects) gives one row
 * of X (the features at the group's centre) and the exact scaled residual in
 * Y (FP64 sum over the cell's bodies minus the softened monopole and
 * quadrupole terms, from the body).  Returns the number of rows the targets
 * yield; only the first cap are written (rows written = min(return, cap)),
 * and which rows those are then depends on the order in which the targets'
 * wavefronts claim them, so a caller that needs every row sizes cap from a
 * first call.  *droppedTargets counts targets whose traversal outgrew its
 * stack; the rows they listed before that are kept.  Fails with "Node array
 * overflow" when the build ran out of nodes. */
long long gd_corrector_samples(const treeState* tree, const float4* pos, const int* targets,
                               int nt, float theta, float minRatio, float eps2, float* X,
                               float* Y, long long cap, int* droppedTargets);
/* gd_walk with quadrupole cells plus the network's residual correction of
 * every accepted cell whose s / r (from the group's centre) exceeds
 * minRatio, evaluated once per (64-body group, cell) on the matrix cores and
 * scaled per body by m s^3 / r^5 (s: the cell size, or its mass radius under
 * the mass-radius opening rule).  Needs a tree built with second and third
 * moments (gd_set_quadrupole(tree, 2)).  phi, when given, receives the
 * quadrupole walk's potential: the network corrects accelerations only. */
void gd_walk_learned(const treeState* tree, const float4* pos, float3* acc, float* phi,
                     float eps2, float G, float theta, int starCount, const gd_mlp* net,
                     float minRatio, gd_walk_stats* stats);

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

/* gd_direct_sum_fp64 without allocating: no counterpart in the reference;
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops).  The partial sums live in a 12 MiB device array that each
 * GPU holds until the process ends; ns may not exceed 524288 (an error
 * otherwise).  Calls that overlap on one GPU share that array and race.
 * Results agree with gd_direct_sum_fp64's (the same kernels; bit for bit
 * whenever both split the sources into the same number of chunks). */
void gd_direct_sum_fp64_noalloc(const float4* pos, const float3* posLo, int starCount,
                                const int* sample, int ns, double eps2, double G, double* acc);

/* Relative force errors against the judge (device arrays):
 * err[k] = |acc[sample[k]] - exact[3k..3k+2]| / |exact[3k..3k+2]| in FP64. */
void gd_force_errors(const float3* acc, const int* sample, int ns, const double* exact,
                     double* err);

/* ---- FP64 diagnostics -----------------------------------------------------
 * out[0..6] (host) = kinetic energy, momentum x/y/z, mass, and the
 * potential energy 0.5 sum m_i phi_i when phi != NULL (else 0), all reduced in
 * FP64 on the device in a fixed order (the same inputs give the same bits);
 * out[6] = total energy.  Synchronises. */
void gd_diagnostics(const float4* pos, const float3* vel, const float* phi, int starCount,
                    double* out);

/* Copies `bytes` (a multiple of 16) from src to dst with a plain streaming
 * kernel (k_stream_copy): a known-size reference for bandwidth and for
 * calibrating hardware byte counters. */
void gd_stream_copy(void* dst, const void* src, size_t bytes);

/* ---- HIP graphs ----------------------------------------------------------------
 * Every launch of the library goes to one stream, created on first use with
 * the default flags and priority.  gd_graph_begin opens a capture of that
 * stream (relaxed mode); the library calls made until gd_graph_end become a
 * graph, instantiated and returned (NULL on failure), and gd_graph_launch
 * replays it on the same stream.  What can be captured: calls that neither
 * synchronise nor allocate -- gd_build_tree_async, gd_walk / gd_walk_learned
 * in the asynchronous mode without statistics or timing, the integrators.
 * Kernels are captured with the device pointers of the moment: a graph of
 * an even number of steps returns the ping-pong buffers (buildTree's pointer
 * swap) to where they were, so it can be replayed step pair after step pair.
 * While a capture is open or a graph lives, the library does not reallocate
 * the buffers graphs address (walk lists, moment and walk arrays): a call
 * that would needs fails instead, and the walk's first pass keeps its list
 * pool (groups without room take the later passes).  A captured walk leaves
 * its status on the device (no deferred report): check gd_tree_summary for an
 * overflowed tree.  Destroy graphs with gd_graph_destroy. */
void gd_graph_begin(void);
void* gd_graph_end(void);
void gd_graph_launch(void* graphExec);
void gd_graph_destroy(void* graphExec);

/* ---- timing ----------------------------------------------------------------
 * HIP events on the library's stream. */
void* gd_event_create(void);
void gd_event_record(void* ev);
float gd_event_elapsed_ms(void* start, void* stop);   /* synchronises on stop */
void gd_event_destroy(void* ev);

#ifdef __cplusplus
}
#endif

#endif
