/*
 * oracle.c -- serial C restatement of the reference's Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the CPU
 * baseline timed by bench.py.  Not part of the product; nothing in the
 * product path links, loads or calls it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, explicit fmaf).
 * Operation order: see oracle.h.
 *
 * Pinning (tests/test_oracle_*.py):
 *   or_bbox            numpy min/max (exact)
 *   or_world_box       known answers + the two float ops recomputed in numpy float32
 *   or_morton          a naive per-bit interleave in numpy + known answers
 *   or_sort_pairs      numpy stable argsort
 *   or_build_topology  an independent recursive octree over sorted keys + invariants
 *   or_tree_mass       FP64 centre of mass (tolerance) + root mass = sum m
 *   or_walk            theta = 0 equals the softened FP32 direct sum up to summation
 *                      order; FP64 direct sum at small theta; a planted-mistake check
 *   or_integrate       closed form in numpy with the same fma order (np.float32 + fma
 *                      emulated in float64, exact for one fma)
 *   or_direct_sum_fp64 momentum conservation (sum m a = 0), two-body closed form
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Bounding box: tree.cu:414-490 (grid-stride partial min/max, then a one-  */
/* block reduction).  Min and max are exact and order-free, so a single     */
/* serial pass gives the same corners.                                      */
/* ------------------------------------------------------------------------ */
void or_bbox(const float4* pos, int n, float3* outMin, float3* outMax) {
    float3 lo = {INFINITY, INFINITY, INFINITY};
    float3 hi = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < n; ++i) {
        lo.x = fminf(lo.x, pos[i].x); hi.x = fmaxf(hi.x, pos[i].x);
        lo.y = fminf(lo.y, pos[i].y); hi.y = fmaxf(hi.y, pos[i].y);
        lo.z = fminf(lo.z, pos[i].z); hi.z = fmaxf(hi.z, pos[i].z);
    }
    *outMin = lo;
    *outMax = hi;
}

/* basicGravity.cu:90-112: extent = max of the three widths, scale =
 * (float)(2^21 - 1) / extent, all in float. */
void or_world_box(const float3* mn, const float3* mx, worldBox* box) {
    float w = mx->x - mn->x;
    float h = mx->y - mn->y;
    float d = mx->z - mn->z;
    float extent = fmaxf(w, fmaxf(h, d));
    float scale = (float)((1u << 21) - 1u) / extent;
    box->minX = mn->x;
    box->minY = mn->y;
    box->minZ = mn->z;
    box->extent = extent;
    box->scale = scale;
}

/* ------------------------------------------------------------------------ */
/* Morton keys: tree.cu:332-367.                                            */
/* ------------------------------------------------------------------------ */

/* float -> u32 as the reference's (uint32_t) cast executes on its GPU
 * (round toward zero, saturating, NaN -> 0).  For the in-range values the
 * key quantisation produces it is the plain C cast. */
static uint32_t f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;          /* NaN, -0, negatives */
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

/* Put bit k of the 21-bit v at bit 3k: a plain loop, one bit at a time. */
static uint64_t or_spread(uint32_t v) {
    uint64_t out = 0;
    v &= 0x1fffffu;                      /* tree.cu:333 keeps 21 bits */
    for (int k = 0; k < 21; ++k) out |= (uint64_t)((v >> k) & 1u) << (3 * k);
    return out;
}

void or_morton(const float4* pos, int n, const worldBox* box, uint64_t* keys, uint32_t* idx) {
    for (int i = 0; i < n; ++i) {
        uint32_t qx = f2u_sat((pos[i].x - box->minX) * box->scale);
        uint32_t qy = f2u_sat((pos[i].y - box->minY) * box->scale);
        uint32_t qz = f2u_sat((pos[i].z - box->minZ) * box->scale);
        /* tree.cu:344: x at bit 3k+2, y at 3k+1, z at 3k */
        keys[i] = (or_spread(qx) << 2) | (or_spread(qy) << 1) | or_spread(qz);
        idx[i] = (uint32_t)i;          /* tree.cu:366 */
    }
}

/* ------------------------------------------------------------------------ */
/* Sort: sorting.cu:65-79 calls a stable LSD radix sort over bits [0,64).  */
/* Restated as a byte-wise stable counting sort, least significant first.   */
/* ------------------------------------------------------------------------ */
void or_sort_pairs(const uint64_t* keysIn, uint64_t* keysOut,
                   const uint32_t* valsIn, uint32_t* valsOut, int n) {
    if (n <= 0) return;
    uint64_t* ka = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    uint64_t* kb = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    uint32_t* va = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t* vb = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    memcpy(ka, keysIn, sizeof(uint64_t) * (size_t)n);
    memcpy(va, valsIn, sizeof(uint32_t) * (size_t)n);
    for (int pass = 0; pass < 8; ++pass) {
        size_t count[257];
        memset(count, 0, sizeof(count));
        int sh = 8 * pass;
        for (int i = 0; i < n; ++i) count[((ka[i] >> sh) & 0xff) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (int i = 0; i < n; ++i) {
            size_t p = count[(ka[i] >> sh) & 0xff]++;
            kb[p] = ka[i];
            vb[p] = va[i];
        }
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    memcpy(keysOut, ka, sizeof(uint64_t) * (size_t)n);
    memcpy(valsOut, va, sizeof(uint32_t) * (size_t)n);
    free(ka); free(kb); free(va); free(vb);
}

/* sorting.cu:99-107 */
void or_reorder(const float4* posIn, const float3* velIn, float4* posOut, float3* velOut,
                const uint32_t* idx, int n) {
    for (int i = 0; i < n; ++i) {
        posOut[i] = posIn[idx[i]];
        velOut[i] = velIn[idx[i]];
    }
}

/* ------------------------------------------------------------------------ */
/* Level-wise topology.                                                     */
/* ------------------------------------------------------------------------ */

/* tree.cu:23-26: the level-L cell of a key is key >> (63 - 3L). */
static uint64_t cell_at(uint64_t key, int level) { return key >> (63 - 3 * level); }

/* tree.cu:50-78, one body: 1 iff the body shares its level-(L-1) cell with a
 * neighbour and begins a new level-L cell. */
static uint32_t start_flag(const uint64_t* k, int i, int n, int level) {
    uint64_t here = cell_at(k[i], level);
    uint64_t up = cell_at(k[i], level - 1);
    int upLeft = (i > 0) && (up == cell_at(k[i - 1], level - 1));
    int upRight = (i < n - 1) && (up == cell_at(k[i + 1], level - 1));
    int sameLeft = (i > 0) && (here == cell_at(k[i - 1], level));
    if (!upLeft && !upRight) return 0u;
    return sameLeft ? 0u : 1u;
}

int or_build_topology(const uint64_t* k, int n, node* nodes, int maxNodes,
                      int* levelOffsets, int* levelCount, int* deepest) {
    /* root, basicGravity.cu:158-169 */
    node* root = &nodes[0];
    root->nodePathFromRoot = 0;
    root->treeLevel = 0;
    root->firstParticleIndex = 0;
    root->particleCount = n;
    root->parentIndex = -1;
    for (int c = 0; c < 8; ++c) root->child[c] = -1;

    uint32_t* flags = (uint32_t*)calloc((size_t)n, sizeof(uint32_t));
    uint32_t* prevFlags = (uint32_t*)calloc((size_t)n, sizeof(uint32_t));
    uint32_t* offs = (uint32_t*)calloc((size_t)n, sizeof(uint32_t));
    uint32_t* prevOffs = (uint32_t*)calloc((size_t)n, sizeof(uint32_t));

    int levelBase = 1, prevBase = 0, rc = 0;
    levelOffsets[0] = 0;
    levelCount[0] = 1;
    *deepest = 0;

    for (int level = 1; level <= 21; ++level) {
        /* identifyNodesAtLevel */
        for (int i = 0; i < n; ++i) flags[i] = start_flag(k, i, n, level);
        /* exclusive scan (tree.cu:286-302) */
        uint32_t run = 0;
        for (int i = 0; i < n; ++i) { offs[i] = run; run += flags[i]; }
        /* createNodes (tree.cu:205-235) */
        for (int i = 0; i < n; ++i) {
            if (!flags[i]) continue;
            uint64_t path = cell_at(k[i], level);
            int idx = (int)offs[i] + levelBase;
            if (idx >= maxNodes) continue;
            node* nd = &nodes[idx];
            nd->nodePathFromRoot = path;
            nd->treeLevel = level;
            nd->firstParticleIndex = i;
            for (int c = 0; c < 8; ++c) nd->child[c] = -1;
            int parent = (level == 1) ? 0 : prevBase + (int)prevOffs[i] + (int)prevFlags[i] - 1;
            nodes[parent].child[path & 7u] = idx;
            nd->parentIndex = parent;
        }
        /* node count, basicGravity.cu:267-282 */
        int count = (n > 0) ? (int)(flags[n - 1] + offs[n - 1]) : 0;
        if (count == 0) {
            levelCount[level] = 0;
            levelOffsets[level] = levelBase;
            break;
        }
        if (count + levelBase > maxNodes) { rc = -1; break; }
        /* updateTreeParticles (tree.cu:181-203) */
        for (int t = 0; t < count; ++t) {
            int idx = levelBase + t;
            int parent = nodes[idx].parentIndex;
            int end;
            if (t + 1 < count && nodes[idx + 1].parentIndex == parent)
                end = nodes[idx + 1].firstParticleIndex;
            else
                end = nodes[parent].firstParticleIndex + nodes[parent].particleCount;
            nodes[idx].particleCount = end - nodes[idx].firstParticleIndex;
        }
        prevBase = levelBase;
        levelOffsets[level] = levelBase;
        levelBase += count;
        levelCount[level] = count;
        uint32_t* t1 = prevOffs; prevOffs = offs; offs = t1;
        uint32_t* t2 = prevFlags; prevFlags = flags; flags = t2;
        *deepest += 1;
    }
    free(flags); free(prevFlags); free(offs); free(prevOffs);
    return rc;
}

/* tree.cu:89-179, one level at a time from the deepest up (basicGravity.cu:317-319). */
int or_tree_mass(node* nodes, const float4* pos, const int* levelOffsets,
                 const int* levelCount, int deepest) {
    int zeroMass = 0;
    for (int level = deepest; level >= 0; --level) {
        for (int t = 0; t < levelCount[level]; ++t) {
            node* nd = &nodes[levelOffsets[level] + t];
            int leaf = 1;
            for (int c = 0; c < 8; ++c) if (nd->child[c] != -1) { leaf = 0; break; }
            float4 out = {0.0f, 0.0f, 0.0f, 0.0f};
            if (leaf) {
                float sx = 0.0f, sy = 0.0f, sz = 0.0f, m = 0.0f;
                for (int b = 0; b < nd->particleCount; ++b) {
                    const float4 p = pos[nd->firstParticleIndex + b];
                    sx = fmaf(p.x, p.w, sx);
                    sy = fmaf(p.y, p.w, sy);
                    sz = fmaf(p.z, p.w, sz);
                    m += p.w;
                }
                if (m > 0.0f) { out.x = sx / m; out.y = sy / m; out.z = sz / m; out.w = m; }
                else zeroMass++;
            } else {
                int kids[8], nk = 0;
                for (int c = 0; c < 8; ++c) if (nd->child[c] != -1) kids[nk++] = nd->child[c];
                float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int c = 0; c < nk; ++c) acc.w += nodes[kids[c]].massData.w;
                for (int c = 0; c < nk; ++c) {
                    const float4 q = nodes[kids[c]].massData;
                    acc.x = fmaf(q.x, q.w, acc.x);
                    acc.y = fmaf(q.y, q.w, acc.y);
                    acc.z = fmaf(q.z, q.w, acc.z);
                }
                if (acc.w > 0.0f) { out.x = acc.x / acc.w; out.y = acc.y / acc.w; out.z = acc.z / acc.w; out.w = acc.w; }
                else zeroMass++;
            }
            nd->massData = out;
        }
    }
    return zeroMass;
}

/* ------------------------------------------------------------------------ */
/* Walk: tree.cu:501-676.  One 32-lane group at a time; the group shares a  */
/* LIFO stack, descends when any valid lane fails the MAC on a non-leaf,    */
/* and otherwise every lane accumulates the node's monopole.                */
/* ------------------------------------------------------------------------ */
#define OR_GROUP 32
#define OR_STACK 168

/* The walk proper.  variant 0 is the reference's rule; variants 1-5 are
 * deliberate mistakes that exist only so the pinning tests can show that
 * their check tells them apart (see or_walk_traced).  When lists != NULL the
 * accepted nodes of every valid body are recorded (at most perBody each,
 * counts in bodyCount, which may exceed perBody). */
static long long walk_core(const float4* pos, int n, const node* nodes, const float3* bmin,
                           const float3* bmax, float eps2, float G, float theta, float3* acc,
                           int g0, int g1, int gstride, long long* opened, int variant,
                           int* bodyCount, int* lists, int perBody) {
    /* tree.cu:506: the x extent only (variant 2: the widest axis instead) */
    float ext = bmax->x - bmin->x;
    if (variant == 2) {
        const float h = bmax->y - bmin->y, d = bmax->z - bmin->z;
        ext = fmaxf(ext, fmaxf(h, d));
    }
    const float th2 = theta * theta;
    const int gsize = variant == 1 ? 1 : OR_GROUP;   /* variant 1: every lane on its own */
    long long accepted = 0, opens = 0;
    if (gstride < 1) gstride = 1;
    if (variant == 1) {   /* the same bodies as groups [g0, g1) of 32 */
        g0 *= OR_GROUP;
        g1 *= OR_GROUP;
        gstride *= OR_GROUP;
    }
    for (int g = g0; g < g1; ++g) {
        if (variant == 1 ? ((g / OR_GROUP) % (gstride / OR_GROUP)) : (g % gstride)) continue;
        float px[OR_GROUP], py[OR_GROUP], pz[OR_GROUP];
        float ax[OR_GROUP], ay[OR_GROUP], az[OR_GROUP];
        int valid[OR_GROUP];
        for (int l = 0; l < gsize; ++l) {
            int i = g * gsize + l;
            valid[l] = i < n;
            px[l] = valid[l] ? pos[i].x : 0.0f;
            py[l] = valid[l] ? pos[i].y : 0.0f;
            pz[l] = valid[l] ? pos[i].z : 0.0f;
            ax[l] = ay[l] = az[l] = 0.0f;
            if (valid[l] && bodyCount) bodyCount[i] = 0;
        }
        int stack[OR_STACK], top = -1;
        for (int c = 0; c < 8; ++c)
            if (nodes[0].child[c] != -1) stack[++top] = nodes[0].child[c];
        while (top >= 0) {
            const int ni = stack[top];
            const node* nd = &nodes[ni];
            const float4 cm = nd->massData;
            float size = (float)((double)ext / pow(2.0, (double)nd->treeLevel));
            int nkids = 0;
            for (int c = 0; c < 8; ++c) nkids += nd->child[c] != -1;
            /* tree.cu:627 (variant 5: without the no-children clause) */
            int leaf = (nd->particleCount <= 1) || (variant != 5 && nkids == 0);
            float dx[OR_GROUP], dy[OR_GROUP], dz[OR_GROUP], d2[OR_GROUP];
            int any = 0, wants = 0;
            for (int l = 0; l < gsize; ++l) {
                dx[l] = px[l] - cm.x;
                dy[l] = py[l] - cm.y;
                dz[l] = pz[l] - cm.z;
                d2[l] = fmaf(dz[l], dz[l], fmaf(dy[l], dy[l], dx[l] * dx[l]));
                int accept = variant == 3 ? (size * size) <= th2 * d2[l]
                                          : (size * size) < th2 * d2[l];
                /* tree.cu:629-634: invalid lanes never vote (variant 4: they do) */
                int want = !(accept || (variant != 4 && !valid[l]) || leaf);
                any |= want;
                wants += want && valid[l];
            }
            --top;
            if (any) {
                opens += wants;
                for (int c = 0; c < 8; ++c)
                    if (nd->child[c] != -1) {
                        if (top + 1 >= OR_STACK) return -1;   /* cannot happen for 21 levels */
                        stack[++top] = nd->child[c];
                    }
                continue;
            }
            for (int l = 0; l < gsize; ++l) {
                float r2 = d2[l] + eps2;
                float mag = -(G * cm.w) / r2;
                float dist = sqrtf(r2);
                ax[l] = fmaf(mag, dx[l] / dist, ax[l]);
                ay[l] = fmaf(mag, dy[l] / dist, ay[l]);
                az[l] = fmaf(mag, dz[l] / dist, az[l]);
                accepted += valid[l];
                if (valid[l] && bodyCount) {
                    const int i = g * gsize + l;
                    if (bodyCount[i] < perBody) lists[(size_t)i * perBody + bodyCount[i]] = ni;
                    bodyCount[i]++;
                }
            }
        }
        for (int l = 0; l < gsize; ++l) {
            if (!valid[l]) continue;
            int i = g * gsize + l;
            acc[i].x = ax[l]; acc[i].y = ay[l]; acc[i].z = az[l];
        }
    }
    if (opened) *opened = opens;
    return accepted;
}

long long or_walk(const float4* pos, int n, const node* nodes, const float3* bmin,
                  const float3* bmax, float eps2, float G, float theta, float3* acc,
                  int g0, int g1, int gstride, long long* opened) {
    return walk_core(pos, n, nodes, bmin, bmax, eps2, G, theta, acc, g0, g1, gstride, opened, 0,
                     NULL, NULL, 0);
}

long long or_walk_traced(const float4* pos, int n, const node* nodes, const float3* bmin,
                         const float3* bmax, float eps2, float G, float theta, float3* acc,
                         long long* opened, int variant, int* bodyCount, int* lists,
                         int perBody) {
    if (variant < 0 || variant > 5) return -1;
    const int groups = (n + OR_GROUP - 1) / OR_GROUP;
    return walk_core(pos, n, nodes, bmin, bmax, eps2, G, theta, acc, 0, groups, 1, opened,
                     variant, bodyCount, lists, perBody);
}

/* tree.cu:678-703: kick with the new acceleration, then drift with the new velocity. */
void or_integrate(float4* pos, float3* vel, const float3* acc, int n, float dt) {
    for (int i = 0; i < n; ++i) {
        float vx = fmaf(acc[i].x, dt, vel[i].x);
        float vy = fmaf(acc[i].y, dt, vel[i].y);
        float vz = fmaf(acc[i].z, dt, vel[i].z);
        vel[i].x = vx; vel[i].y = vy; vel[i].z = vz;
        pos[i].x = fmaf(vx, dt, pos[i].x);
        pos[i].y = fmaf(vy, dt, pos[i].y);
        pos[i].z = fmaf(vz, dt, pos[i].z);
    }
}

void or_direct_sum_fp64(const float4* pos, int n, const int* sample, int ns,
                        double eps2, double G, double* acc) {
    for (int k = 0; k < ns; ++k) {
        int i = sample[k];
        double xi = pos[i].x, yi = pos[i].y, zi = pos[i].z;
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx = xi - (double)pos[j].x;
            double dy = yi - (double)pos[j].y;
            double dz = zi - (double)pos[j].z;
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv = 1.0 / sqrt(r2);
            double f = -G * (double)pos[j].w * inv * inv * inv;
            ax += f * dx; ay += f * dy; az += f * dz;
        }
        acc[3 * k] = ax; acc[3 * k + 1] = ay; acc[3 * k + 2] = az;
    }
}

/* basicGravity.cu:116-325 then :779-783 on host arrays. */
int or_step(float4** pos, float3** vel, float4** posSpare, float3** velSpare, float3* acc,
            int n, uint64_t* keysIn, uint64_t* keysOut, uint32_t* idxIn, uint32_t* idxOut,
            node* nodes, int maxNodes, int* levelOffsets, int* levelCount, int* deepest,
            float eps2, float G, float theta, float dt) {
    float3 mn, mx;
    worldBox box;
    or_bbox(*pos, n, &mn, &mx);
    or_world_box(&mn, &mx, &box);
    or_morton(*pos, n, &box, keysIn, idxIn);
    or_sort_pairs(keysIn, keysOut, idxIn, idxOut, n);
    or_reorder(*pos, *vel, *posSpare, *velSpare, idxOut, n);
    float4* tp = *pos; *pos = *posSpare; *posSpare = tp;       /* basicGravity.cu:220-226 */
    float3* tv = *vel; *vel = *velSpare; *velSpare = tv;
    if (or_build_topology(keysOut, n, nodes, maxNodes, levelOffsets, levelCount, deepest))
        return -1;
    or_tree_mass(nodes, *pos, levelOffsets, levelCount, *deepest);
    int groups = (n + OR_GROUP - 1) / OR_GROUP;
    or_walk(*pos, n, nodes, &mn, &mx, eps2, G, theta, acc, 0, groups, 1, NULL);
    or_integrate(*pos, *vel, acc, n, dt);
    return 0;
}
