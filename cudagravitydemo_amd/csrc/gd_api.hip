/*
 * gd_api.hip -- the C host API: errors, device memory, sorter / builder
 * handles, initialiseTree, createWorldBox, buildTree and the step pieces
 * that have no reference counterpart.
 *
 * Reference host code restated:
 *   sorterCreate / radixSortPairs / sorterDestroyer   sorting.cu:15-92
 *   sumCreate / prefixSum / sumDestroyer              tree.cu:241-318
 *   initialiseTree                                    basicGravity.cu:5-77
 *   createWorldBox                                    basicGravity.cu:79-113
 *   buildTree                                         basicGravity.cu:116-325
 *   roundNumberToMultiple / blockNumCalculator        basicGravity.cu:351-427
 *   G, theta                                          units.c:3-4
 * Every launch goes to one library stream (gd_stream), which synchronises
 * with the null stream the reference used, so synchronous copies order after
 * it as in the reference.
 */
#include "gd_internal.h"
#include "gravity_host.h"

#include <stdarg.h>
#include <string.h>
#include <vector>
#include <sys/sysinfo.h>

extern "C" {
float G = 1.0f;
float theta = 0.7f;
}

/* ---------------------------------------------------------------------------
 * Errors
 * ------------------------------------------------------------------------- */
static int g_exitOnError = 1;
static char g_lastError[512] = "";

void gd_fail(const char* file, int line, const char* fmt, ...) {
    char msg[400];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    if (g_exitOnError) {
        printf("\nERROR at: %s:%d\n%s\n", file, line, msg);
        fflush(stdout);
        exit(-1);
    }
    if (g_lastError[0] == '\0') snprintf(g_lastError, sizeof(g_lastError), "%s:%d: %s", file, line, msg);
}

int gd_has_error(void) { return g_lastError[0] != '\0'; }

/* ---------------------------------------------------------------------------
 * The library's stream: every kernel, asynchronous copy and event of the
 * library goes to this one stream, created on first use with the default
 * flags (it synchronises with the null stream, so the synchronous copies of
 * gd_memcpy_* and hipMemcpy order after it) and the default priority.
 * ------------------------------------------------------------------------- */
static hipStream_t g_stream = nullptr;
static int g_capturing = 0;      /* between gd_graph_begin and gd_graph_end */
static int g_liveGraphs = 0;     /* instantiated graphs not destroyed yet */

hipStream_t gd_stream(void) {
    if (!g_stream) GD_HIP(hipStreamCreate(&g_stream));
    return g_stream;
}

/* A second stream, also with the default flags and priority, for the few
 * groups a synchronous walk sends past its first pass: their latency-bound
 * kernels run beside the first pass's evaluation, and the library stream
 * waits for them (gd_walk.hip). */
static hipStream_t g_sideStream = nullptr;
hipStream_t gd_side_stream(void) {
    if (!g_sideStream) GD_HIP(hipStreamCreate(&g_sideStream));
    return g_sideStream;
}
int gd_capturing(void) { return g_capturing; }
int gd_graphs_hold_buffers(void) { return g_capturing || g_liveGraphs > 0; }

extern "C" {

void gd_set_exit_on_error(int exitOnError) { g_exitOnError = exitOnError ? 1 : 0; }
const char* gd_last_error(void) { return g_lastError; }
void gd_clear_error(void) {
    g_lastError[0] = '\0';
    (void)hipGetLastError();
}

/* ---------------------------------------------------------------------------
 * Device memory
 * ------------------------------------------------------------------------- */
void* gd_malloc(size_t bytes) {
    void* p = nullptr;
    GD_HIP(hipMalloc(&p, bytes ? bytes : 1));
    return p;
}
void gd_free(void* p) {
    if (p) GD_HIP(hipFree(p));
}
void gd_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (bytes) GD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
void gd_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (bytes) GD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
void gd_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    if (bytes) GD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
}
void gd_memset(void* p, int value, size_t bytes) {
    if (bytes) GD_HIP(hipMemset(p, value, bytes));
}
void gd_synchronize(void) { GD_HIP(hipDeviceSynchronize()); }
int gd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
void gd_device_name(char* buf, int bufLen) {
    hipDeviceProp_t prop;
    int dev = 0;
    GD_HIP(hipGetDevice(&dev));
    GD_HIP(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, (size_t)bufLen, "%s %s CUs=%d", prop.gcnArchName, prop.name,
             prop.multiProcessorCount);
}

/* ---------------------------------------------------------------------------
 * Sorter (sorting.cu:15-92)
 * ------------------------------------------------------------------------- */
radixSorter* sorterCreate(int maxCount) {
    if (maxCount <= 0) {
        GD_FAIL("Attempted to create a radix sort over an array of <= 0 elements.");
        return nullptr;
    }
    gd_sorter* s = (gd_sorter*)calloc(1, sizeof(gd_sorter));
    if (!s) {
        GD_FAIL("Radix sorter creation failed.");
        return nullptr;
    }
    s->maxTiles = gd_div_up(maxCount, GD_SORT_TILE);
    const size_t statusBytes = sizeof(unsigned long long) * (size_t)s->maxTiles * GD_SORT_BINS;
    GD_HIP(hipMalloc(&s->keysTmp, sizeof(uint64_t) * (size_t)maxCount));
    GD_HIP(hipMalloc(&s->valsTmp, sizeof(uint32_t) * (size_t)maxCount));
    GD_HIP(hipMalloc(&s->hist, sizeof(uint32_t) * GD_SORT_PASSES * GD_SORT_GROUPS * GD_SORT_BINS));
    GD_HIP(hipMalloc(&s->tileCounter, sizeof(uint32_t) * GD_SORT_PASSES * GD_SORT_GROUPS));
    GD_HIP(hipMalloc(&s->status, statusBytes));
    GD_HIP(hipMemset(s->status, 0, statusBytes));
    GD_HIP(hipMalloc(&s->epoch, sizeof(uint32_t)));
    GD_HIP(hipMemset(s->epoch, 0, sizeof(uint32_t)));
    s->pub.tempStorage = s->status;
    s->pub.tempStorageBytes = statusBytes +
                              sizeof(uint32_t) * GD_SORT_PASSES * GD_SORT_GROUPS *
                                  (GD_SORT_BINS + 1) +
                              (sizeof(uint64_t) + sizeof(uint32_t)) * (size_t)maxCount;
    s->pub.capacity = maxCount;
    if (gd_has_error()) {
        free(s);
        return nullptr;
    }
    return &s->pub;
}

void radixSortPairs(radixSorter* sorter, const uint64_t* keysIn, uint64_t* keysOut,
                    const uint32_t* valsIn, uint32_t* valsOut, int numCount) {
    if (sorter == nullptr) {
        GD_FAIL("Sorter struct passed to radix sort was uninitialised");
        return;
    }
    if (numCount <= 0 || numCount > sorter->capacity) {
        GD_FAIL("Length of array to sort conflicts with capacity of radix sorter.");
        return;
    }
    gd_sort_pairs(gd_sorter_of(sorter), keysIn, keysOut, valsIn, valsOut, numCount);
}

void sorterDestroyer(radixSorter* sorter) {
    if (sorter == nullptr) {
        GD_FAIL("Attempted to delete non-existant sorter.");
        return;
    }
    gd_sorter* s = gd_sorter_of(sorter);
    GD_HIP(hipFree(s->keysTmp));
    GD_HIP(hipFree(s->valsTmp));
    GD_HIP(hipFree(s->hist));
    GD_HIP(hipFree(s->tileCounter));
    GD_HIP(hipFree(s->status));
    GD_HIP(hipFree(s->epoch));
    memset(s, 0, sizeof(*s));
    free(s);
}

/* ---------------------------------------------------------------------------
 * Builder (tree.cu:241-318)
 * ------------------------------------------------------------------------- */
static void gd_builder_walk_arrays(gd_builder* b, int maxNodes) {
    if (maxNodes <= b->maxNodes) return;
    if (gd_graphs_hold_buffers()) {
        GD_FAIL("the walk arrays cannot be allocated while a graph holds the library's "
                "buffers (build once before capturing)");
        return;
    }
    if (b->wcom) GD_HIP(hipFree(b->wcom));
    if (b->wlink) GD_HIP(hipFree(b->wlink));
    if (b->wrad) GD_HIP(hipFree(b->wrad));
    if (b->meta) GD_HIP(hipFree(b->meta));
    GD_HIP(hipMalloc(&b->wcom, sizeof(float4) * (size_t)maxNodes));
    GD_HIP(hipMalloc(&b->wlink, sizeof(int2) * (size_t)maxNodes));
    GD_HIP(hipMalloc(&b->wrad, sizeof(float) * (size_t)maxNodes));
    GD_HIP(hipMalloc(&b->meta, sizeof(int4) * (size_t)maxNodes));
    b->maxNodes = maxNodes;
}

treeBuilder* sumCreate(int maxCount) {
    if (maxCount <= 0) {
        GD_FAIL("Attempted to initialise a prefix sum over an array of <= 0 elements.");
        return nullptr;
    }
    gd_builder* b = (gd_builder*)calloc(1, sizeof(gd_builder));
    if (!b) {
        GD_FAIL("Prefix sum could not be initialised.");
        return nullptr;
    }
    const size_t arr = sizeof(uint32_t) * (size_t)maxCount;
    GD_HIP(hipMalloc(&b->pub.flagsPing, arr));
    GD_HIP(hipMalloc(&b->pub.flagsPong, arr));
    GD_HIP(hipMalloc(&b->pub.offsetsPing, arr));
    GD_HIP(hipMalloc(&b->pub.offsetsPong, arr));
    b->maxTiles = gd_div_up(maxCount, GD_BUILD_TILE);
    const size_t tiles = (size_t)GD_LEVELS * b->maxTiles;
    GD_HIP(hipMalloc(&b->tileCounts, sizeof(uint32_t) * tiles));
    GD_HIP(hipMalloc(&b->tileOffsets, sizeof(uint32_t) * tiles));
    GD_HIP(hipMalloc(&b->tileFirst, sizeof(int) * tiles));
    GD_HIP(hipMalloc(&b->tileNext, sizeof(int) * tiles));
    GD_HIP(hipMalloc(&b->levelTotals, sizeof(uint32_t) * GD_LEVELS));
    GD_HIP(hipMemset(b->levelTotals, 0, sizeof(uint32_t) * GD_LEVELS));
    GD_HIP(hipMalloc(&b->dinfo, sizeof(gd_tree_info)));
    GD_HIP(hipMemset(b->dinfo, 0, sizeof(gd_tree_info)));
    GD_HIP(hipHostMalloc(&b->hinfo, sizeof(gd_tree_info), hipHostMallocDefault));
    GD_HIP(hipMalloc(&b->bboxPartMin, sizeof(float3) * GD_BBOX_BLOCKS));
    GD_HIP(hipMalloc(&b->bboxPartMax, sizeof(float3) * GD_BBOX_BLOCKS));
    GD_HIP(hipMalloc(&b->bboxPartCom, 4 * sizeof(double) * GD_BBOX_BLOCKS));
    GD_HIP(hipMalloc(&b->scanTiles, sizeof(uint32_t) * (size_t)gd_div_up(maxCount, GD_SCAN_TILE)));
    b->pub.tempStorage = b->scanTiles;
    b->pub.tempStorageBytes = sizeof(uint32_t) * (size_t)gd_div_up(maxCount, GD_SCAN_TILE);
    b->pub.capacity = maxCount;
    if (gd_has_error()) {
        free(b);
        return nullptr;
    }
    return &b->pub;
}

void prefixSum(treeBuilder* builder, uint32_t* flags, uint32_t* offsets, int maxCount) {
    if (builder == nullptr) {
        GD_FAIL("treeBuilder struct passed to prefixSum() was uninitialised.");
        return;
    }
    if (maxCount <= 0 || maxCount > builder->capacity) {
        GD_FAIL("Length of array to sum conflicts with capacity of treeBuilder struct.");
        return;
    }
    gd_exclusive_scan_u32(gd_builder_of(builder), flags, offsets, maxCount);
}

void sumDestroyer(treeBuilder* sum) {
    if (sum == nullptr) {
        GD_FAIL("Attempted to delete non-existant prefix sum object.");
        return;
    }
    gd_builder* b = gd_builder_of(sum);
    GD_HIP(hipFree(b->pub.flagsPing));
    GD_HIP(hipFree(b->pub.flagsPong));
    GD_HIP(hipFree(b->pub.offsetsPing));
    GD_HIP(hipFree(b->pub.offsetsPong));
    GD_HIP(hipFree(b->tileCounts));
    GD_HIP(hipFree(b->tileOffsets));
    GD_HIP(hipFree(b->tileFirst));
    GD_HIP(hipFree(b->tileNext));
    GD_HIP(hipFree(b->levelTotals));
    GD_HIP(hipFree(b->dinfo));
    GD_HIP(hipHostFree(b->hinfo));
    GD_HIP(hipFree(b->bboxPartMin));
    GD_HIP(hipFree(b->bboxPartMax));
    GD_HIP(hipFree(b->bboxPartCom));
    GD_HIP(hipFree(b->scanTiles));
    if (b->wcom) GD_HIP(hipFree(b->wcom));
    if (b->wlink) GD_HIP(hipFree(b->wlink));
    if (b->wrad) GD_HIP(hipFree(b->wrad));
    if (b->meta) GD_HIP(hipFree(b->meta));
    if (b->wmom) GD_HIP(hipFree(b->wmom));
    memset(b, 0, sizeof(*b));
    free(b);
}

/* ---------------------------------------------------------------------------
 * Host orchestration (basicGravity.cu)
 * ------------------------------------------------------------------------- */
void initialiseTree(treeState* state, int starCount, int paddedStarCount, int integrateBlocks) {
    if (!state) {
        GD_FAIL("initialiseTree: NULL treeState");
        return;
    }
    if (paddedStarCount < starCount) paddedStarCount = starCount;
    radixSorter* sorter = sorterCreate(starCount);
    treeBuilder* builder = sumCreate(starCount);
    if (!sorter || !builder) return;
    const long long maxNodes = 8ll * starCount;   /* basicGravity.cu:11 */
    if (maxNodes > 0x7fffffffll) {
        GD_FAIL("initialiseTree: %d bodies need more than 2^31 nodes", starCount);
        return;
    }
    state->sorter = sorter;
    state->builder = builder;
    state->starCount = starCount;
    state->maxNodes = (int)maxNodes;
    state->treeBlocks = integrateBlocks;
    GD_HIP(hipMalloc(&state->nodes, sizeof(node) * (size_t)maxNodes));
    GD_HIP(hipMalloc(&state->mortonIn, sizeof(uint64_t) * (size_t)starCount));
    GD_HIP(hipMalloc(&state->mortonOut, sizeof(uint64_t) * (size_t)starCount));
    GD_HIP(hipMalloc(&state->indexIn, sizeof(uint32_t) * (size_t)starCount));
    GD_HIP(hipMalloc(&state->indexOut, sizeof(uint32_t) * (size_t)starCount));
    GD_HIP(hipMalloc(&state->sortedPosMassVals, sizeof(float4) * (size_t)paddedStarCount));
    GD_HIP(hipMalloc(&state->sortedVelVals, sizeof(float3) * (size_t)paddedStarCount));
    GD_HIP(hipMemset(state->sortedPosMassVals, 0, sizeof(float4) * (size_t)paddedStarCount));
    GD_HIP(hipMemset(state->sortedVelVals, 0, sizeof(float3) * (size_t)paddedStarCount));
    /* indexIn = iota (basicGravity.cu:58-64) */
    uint32_t* iota = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)starCount);
    for (int i = 0; i < starCount; ++i) iota[i] = (uint32_t)i;
    GD_HIP(hipMemcpy(state->indexIn, iota, sizeof(uint32_t) * (size_t)starCount,
                     hipMemcpyHostToDevice));
    free(iota);
    GD_HIP(hipMalloc(&state->boundingData[0], sizeof(float3) * BOXBLOCKS));
    GD_HIP(hipMalloc(&state->boundingData[1], sizeof(float3) * BOXBLOCKS));
    GD_HIP(hipMalloc(&state->boundingData[2], sizeof(float3)));
    GD_HIP(hipMalloc(&state->boundingData[3], sizeof(float3)));
    gd_builder_walk_arrays(gd_builder_of(builder), (int)maxNodes);
    state->deepestLevel = 0;
}

void createWorldBox(const float4* cudaPositionMassVals, const int starCount,
                    float3* cudaMinCorner, float3* cudaMaxCorner, float3* cudaOutputMin,
                    float3* cudaOutputMax, worldBox* outputWorldBox) {
    float3 mn, mx;
    localMinMaxFinder(cudaPositionMassVals, (unsigned)starCount, cudaMinCorner, cudaMaxCorner);
    globalMinMaxReducer(cudaMinCorner, cudaMaxCorner, BOXBLOCKS, cudaOutputMin, cudaOutputMax);
    GD_HIP(hipMemcpy(&mn, cudaOutputMin, sizeof(float3), hipMemcpyDeviceToHost));
    GD_HIP(hipMemcpy(&mx, cudaOutputMax, sizeof(float3), hipMemcpyDeviceToHost));
    const float w = mx.x - mn.x, h = mx.y - mn.y, d = mx.z - mn.z;
    const float extent = fmaxf(w, fmaxf(h, d));
    const int biggest = (1 << 21) - 1;
    const float scale = (float)biggest / extent;
    outputWorldBox->extent = extent;
    outputWorldBox->minX = mn.x;
    outputWorldBox->minY = mn.y;
    outputWorldBox->minZ = mn.z;
    outputWorldBox->scale = scale;
}

/* optional HIP-event timing of the build's stages (library stream) */
static int g_buildTiming = 0;
static int g_buildTimed = 0;
static hipEvent_t g_buildEv[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

static void gd_build_mark(int k) {
    if (g_buildTiming) GD_HIP(hipEventRecord(g_buildEv[k], gd_stream()));
}

void gd_build_timing(int enable) {
    g_buildTiming = enable ? 1 : 0;
    if (g_buildTiming && !g_buildEv[0])
        for (int e = 0; e < 7; ++e) GD_HIP(hipEventCreate(&g_buildEv[e]));
}

int gd_build_phase_ms(float* ms) {
    if (!g_buildTimed) return -1;
    GD_HIP(hipEventSynchronize(g_buildEv[6]));
    for (int e = 0; e < 6; ++e) GD_HIP(hipEventElapsedTime(&ms[e], g_buildEv[e], g_buildEv[e + 1]));
    return 0;
}

static void gd_build_tree(treeState* tree, float4** pos, float3** vel, float3** posLo,
                          float3** spareLo, int records) {
    if (!tree || !tree->sorter || !tree->builder || !pos || !vel) {
        GD_FAIL("buildTree: uninitialised tree state");
        return;
    }
    const int n = tree->starCount;
    if (n <= 0) return;
    gd_builder* b = gd_builder_of(tree->builder);
    if (gd_capturing() && g_buildTiming) {
        GD_FAIL("buildTree: a captured build cannot be timed (gd_build_timing(0) first)");
        return;
    }
    gd_builder_walk_arrays(b, tree->maxNodes);
    if (gd_has_error()) return;
    gd_build_mark(0);
    /* 1. bbox and quantisation box, on the device */
    gd_launch_bbox(*pos, (posLo && *posLo) ? *posLo : nullptr, n, b, tree->boundingData[2],
                   tree->boundingData[3]);
    gd_build_mark(1);
    /* 2. keys */
    gd_launch_morton_dev(*pos, n, b->dinfo, tree->mortonIn, tree->indexIn);
    gd_build_mark(2);
    /* 3. sort */
    gd_sort_pairs(gd_sorter_of(tree->sorter), tree->mortonIn, tree->mortonOut, tree->indexIn,
                  tree->indexOut, n);
    gd_build_mark(3);
    /* 4. reorder into the spare buffers, then swap (basicGravity.cu:217-226) */
    if (posLo && spareLo && *posLo && *spareLo) {
        gd_launch_reorder_lo(*pos, *vel, *posLo, tree->sortedPosMassVals, tree->sortedVelVals,
                             *spareLo, tree->indexOut, n);
        float3* t = *posLo;
        *posLo = *spareLo;
        *spareLo = t;
    } else {
        gd_launch_reorder(*pos, *vel, tree->sortedPosMassVals, tree->sortedVelVals,
                          tree->indexOut, n);
    }
    float4* tp = *pos;
    *pos = tree->sortedPosMassVals;
    tree->sortedPosMassVals = tp;
    float3* tv = *vel;
    *vel = tree->sortedVelVals;
    tree->sortedVelVals = tv;
    gd_build_mark(4);
    /* 5. topology and centre of mass */
    gd_build_mid_event = g_buildTiming ? g_buildEv[5] : nullptr;
    gd_build_octree(b, tree->mortonOut, *pos, n, tree->nodes, tree->maxNodes, records);
    gd_build_mid_event = nullptr;
    gd_build_mark(6);
    if (g_buildTiming) g_buildTimed = 1;
}

void gd_build_tree_async(treeState* tree, float4** pos, float3** vel, float3** posLo,
                         float3** spareLo) {
    const int records = tree && tree->builder ? !gd_builder_of(tree->builder)->noRecords : 1;
    gd_build_tree(tree, pos, vel, posLo, spareLo, records);
}

void gd_set_node_records(treeState* tree, int enable) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_set_node_records: uninitialised tree state");
        return;
    }
    gd_builder_of(tree->builder)->noRecords = enable ? 0 : 1;
}

void gd_write_node_records(treeState* tree) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_write_node_records: uninitialised tree state");
        return;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    if (b->lastBodies <= 0) {
        GD_FAIL("gd_write_node_records: no tree has been built");
        return;
    }
    gd_build_node_records(b, tree->mortonOut, tree->nodes, tree->maxNodes);
}

void gd_set_quadrupole(treeState* tree, int enable) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_set_quadrupole: uninitialised tree state");
        return;
    }
    if (enable < 0 || enable > 2) {
        GD_FAIL("gd_set_quadrupole: enable must be 0, 1 or 2 (got %d)", enable);
        return;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    if (gd_graphs_hold_buffers() && enable && b->momNodes < tree->maxNodes) {
        GD_FAIL("gd_set_quadrupole: the moment arrays cannot be allocated while a graph holds "
                "the library's buffers");
        return;
    }
    b->quad = enable;
    if (b->quad && b->momNodes < tree->maxNodes) {
        if (b->wmom) GD_HIP(hipFree(b->wmom));
        GD_HIP(hipMalloc(&b->wmom, 4 * sizeof(float4) * (size_t)tree->maxNodes));
        b->momNodes = tree->maxNodes;
    }
}

/* The moment records of nodes [0, count) in host memory; the build does not
 * store the (zero) moments of single-body leaves, which read as zero here. */
static std::vector<float4> gd_moment_records(const gd_builder* b, int count) {
    std::vector<float4> rec((size_t)4 * (count > 0 ? count : 0));
    if (count <= 0 || gd_has_error()) return rec;
    GD_HIP(hipMemcpy(rec.data(), b->wmom, 4 * sizeof(float4) * (size_t)count,
                     hipMemcpyDeviceToHost));
    std::vector<int2> link((size_t)count);
    GD_HIP(hipMemcpy(link.data(), b->wlink, sizeof(int2) * (size_t)count, hipMemcpyDeviceToHost));
    if (gd_has_error()) return rec;
    for (int i = 0; i < count; ++i)
        if (link[i].x < 0 && (link[i].y >> 5) == 1)
            for (int k = 0; k < 4; ++k) rec[(size_t)4 * i + k] = make_float4(0.f, 0.f, 0.f, 0.f);
    return rec;
}

void gd_tree_third_moments(const treeState* tree, void* oct, int count) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_tree_third_moments: uninitialised tree state");
        return;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    if (count < 0 || count > b->maxNodes) {
        GD_FAIL("gd_tree_third_moments: count %d outside [0, %d]", count, b->maxNodes);
        return;
    }
    if (b->quad != 2 || !b->wmom) {
        GD_FAIL("gd_tree_third_moments: third moments are off (gd_set_quadrupole(tree, 2))");
        return;
    }
    const std::vector<float4> rec = gd_moment_records(b, count);
    if (gd_has_error()) return;
    float4* out = static_cast<float4*>(oct);
    for (int i = 0; i < count; ++i) {
        const float4 m1 = rec[4 * (size_t)i + 1], m2 = rec[4 * (size_t)i + 2],
                     m3 = rec[4 * (size_t)i + 3];
        out[3 * (size_t)i] = make_float4(m1.z, m1.w, m2.x, m2.y);
        out[3 * (size_t)i + 1] = make_float4(m2.z, m2.w, m3.x, m3.y);
        out[3 * (size_t)i + 2] = make_float4(m3.z, m3.w, 0.f, 0.f);
    }
}

void gd_tree_walk_arrays(const treeState* tree, void* com, void* link, void* quad, int count) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_tree_walk_arrays: uninitialised tree state");
        return;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    if (count < 0 || count > b->maxNodes) {
        GD_FAIL("gd_tree_walk_arrays: count %d outside [0, %d]", count, b->maxNodes);
        return;
    }
    if (com) GD_HIP(hipMemcpy(com, b->wcom, sizeof(float4) * (size_t)count, hipMemcpyDeviceToHost));
    if (link) GD_HIP(hipMemcpy(link, b->wlink, sizeof(int2) * (size_t)count, hipMemcpyDeviceToHost));
    if (quad && b->quad && b->wmom) {
        const std::vector<float4> rec = gd_moment_records(b, count);
        if (gd_has_error()) return;
        float4* out = static_cast<float4*>(quad);
        for (int i = 0; i < count; ++i) {
            const float4 m0 = rec[4 * (size_t)i], m1 = rec[4 * (size_t)i + 1];
            const float tr = m0.x + m0.y + m0.z;   /* as the walks form it */
            out[2 * (size_t)i] = m0;
            out[2 * (size_t)i + 1] = make_float4(m1.x, m1.y, tr, 0.f);
        }
    }
}

void gd_tree_mass_radii(const treeState* tree, float* rad, int count) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_tree_mass_radii: uninitialised tree state");
        return;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    if (count < 0 || count > b->maxNodes) {
        GD_FAIL("gd_tree_mass_radii: count %d outside [0, %d]", count, b->maxNodes);
        return;
    }
    if (rad) GD_HIP(hipMemcpy(rad, b->wrad, sizeof(float) * (size_t)count, hipMemcpyDeviceToHost));
}

void gd_tree_levels(treeState* tree) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_tree_levels: uninitialised tree state");
        return;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    GD_HIP(hipMemcpy(b->hinfo, b->dinfo, sizeof(gd_tree_info), hipMemcpyDeviceToHost));
    if (gd_has_error()) return;
    const gd_tree_info* h = b->hinfo;
    if (h->box.scale == 0.0f) {
        GD_FAIL("rootBox failed to initialise.");
        return;
    }
    if (h->overflow) {
        int level = 1, run = 1;
        for (; level <= GD_MAX_LEVEL; ++level) {
            run += h->levelCount[level];
            if (run > tree->maxNodes) break;
        }
        GD_FAIL("Node array overflow at level: %d", level);
        return;
    }
    tree->deepestLevel = h->deepestLevel;
    /* levels 0..deepest, plus the empty level that ended the reference's loop */
    const int last = h->deepestLevel < GD_MAX_LEVEL ? h->deepestLevel + 1 : GD_MAX_LEVEL;
    for (int L = 0; L <= last; ++L) {
        tree->levelOffsets[L] = h->levelOffsets[L];
        tree->levelCount[L] = h->levelCount[L];
    }
}

void buildTree(treeState* tree, float4** cudaPosMasVals, float3** cudaVelVals) {
    gd_build_tree(tree, cudaPosMasVals, cudaVelVals, nullptr, nullptr, 1);
    if (gd_has_error()) return;
    gd_tree_levels(tree);
}

void gd_tree_root_fp64(const treeState* tree, double* out) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_tree_root_fp64: uninitialised tree state");
        return;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    GD_HIP(hipMemcpy(out, b->dinfo->rootCom, 4 * sizeof(double), hipMemcpyDeviceToHost));
}

void gd_tree_summary(const treeState* tree, int* totalNodes, int* overflow, int* zeroMassNodes) {
    if (!tree || !tree->builder) {
        GD_FAIL("gd_tree_summary: uninitialised tree state");
        return;
    }
    gd_builder* b = gd_builder_of(tree->builder);
    GD_HIP(hipMemcpy(b->hinfo, b->dinfo, sizeof(gd_tree_info), hipMemcpyDeviceToHost));
    if (totalNodes) *totalNodes = b->hinfo->totalNodes;
    if (overflow) *overflow = b->hinfo->overflow;
    if (zeroMassNodes) *zeroMassNodes = b->hinfo->zeroMassNodes;
}

int roundNumberToMultiple(int number, int multiple) {
    if (multiple == 0) return number;
    if (multiple < 0) {
        GD_FAIL("Rounding up to a negative number is undefined");
        return number;
    }
    const int r = number % multiple;
    return r == 0 ? number : number + multiple - r;
}

void blockNumCalculator(int* accelerationBlocks, int* treeBuilderBlocks, int* integrateBlocks,
                        int* paddedStarCount, int* integrationThreadsPerBlock,
                        int* treeBuilderThreadsPerBlock, int starCount) {
    const int accelThreads = WARPTHREADS;
    *integrationThreadsPerBlock = 256;
    *treeBuilderThreadsPerBlock = 256;
    printf("CPU CORES: %d\n", get_nprocs());
    *paddedStarCount = roundNumberToMultiple(starCount, accelThreads);
    *accelerationBlocks = *paddedStarCount / accelThreads + (starCount % accelThreads ? 1 : 0);
    *integrateBlocks = *paddedStarCount / 256 + (starCount % 256 ? 1 : 0);
    *treeBuilderBlocks = *paddedStarCount / 256 + (starCount % 256 ? 1 : 0);
}

/* ---------------------------------------------------------------------------
 * HIP graphs of the library's work
 * ------------------------------------------------------------------------- */
void gd_graph_begin(void) {
    if (g_capturing) {
        GD_FAIL("gd_graph_begin: a capture is already open");
        return;
    }
    GD_HIP(hipStreamBeginCapture(gd_stream(), hipStreamCaptureModeRelaxed));
    if (!gd_has_error()) g_capturing = 1;
}

void* gd_graph_end(void) {
    if (!g_capturing) {
        GD_FAIL("gd_graph_end: no capture is open");
        return nullptr;
    }
    hipGraph_t graph = nullptr;
    const hipError_t err = hipStreamEndCapture(gd_stream(), &graph);
    g_capturing = 0;
    if (err != hipSuccess || !graph) {
        GD_FAIL("gd_graph_end: capture failed: %s", hipGetErrorString(err));
        return nullptr;
    }
    if (gd_has_error()) {   /* a call inside the capture failed: no graph */
        GD_HIP(hipGraphDestroy(graph));
        return nullptr;
    }
    hipGraphExec_t exec = nullptr;
    GD_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    GD_HIP(hipGraphDestroy(graph));
    if (gd_has_error()) return nullptr;
    ++g_liveGraphs;
    return (void*)exec;
}

void gd_graph_launch(void* exec) {
    if (!exec) {
        GD_FAIL("gd_graph_launch: no graph");
        return;
    }
    GD_HIP(hipGraphLaunch((hipGraphExec_t)exec, gd_stream()));
}

void gd_graph_destroy(void* exec) {
    if (!exec) return;
    GD_HIP(hipGraphExecDestroy((hipGraphExec_t)exec));
    if (g_liveGraphs > 0) --g_liveGraphs;
}

/* ---------------------------------------------------------------------------
 * Events
 * ------------------------------------------------------------------------- */
void* gd_event_create(void) {
    hipEvent_t e = nullptr;
    GD_HIP(hipEventCreate(&e));
    return (void*)e;
}
void gd_event_record(void* ev) { GD_HIP(hipEventRecord((hipEvent_t)ev, gd_stream())); }
float gd_event_elapsed_ms(void* start, void* stop) {
    float ms = 0.f;
    GD_HIP(hipEventSynchronize((hipEvent_t)stop));
    GD_HIP(hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)stop));
    return ms;
}
void gd_event_destroy(void* ev) {
    if (ev) GD_HIP(hipEventDestroy((hipEvent_t)ev));
}

}  /* extern "C" */
