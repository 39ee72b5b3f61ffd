/*
 * gravity_demo.c -- the reference's main loop (basicGravity.cu:458-837) as a
 * plain C program against libgravity.so.  It shows the drop-in boundary: the
 * loop is the reference's, with each <<<grid, block>>> launch replaced by the
 * launcher of the same name and cudaMalloc / cudaMemcpy by gd_malloc /
 * gd_memcpy_*.  No HIP header is included.
 *
 * Usage: gravity_demo [N] [steps] [theta] [eps] [dt] [frames_every] [outfile] [icfile]
 * icfile (optional) holds the initial conditions instead of the built-in disk:
 * int32 N, float32 (x, y, z, m) x N, float32 (vx, vy, vz) x N; N then comes
 * from the file.
 * Writes the reference's trajectory format (basicGravity.cu:614-635,793-801):
 * int32 N, int32 frames, RGB[N], then float32[3N] per frame, bodies in
 * original order (the original index of every sorted slot is carried
 * through each step's permutation).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gravity_ext.h"
#include "gravity_host.h"

/* splitmix64 (Steele, Lea & Flood 2014) for a seeded uniform in [0, 1) */
static uint64_t g_state = 0x2545F4914F6CDD1Dull;
static double uniform01(void) {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* One exponential disk (scale length 1, sech^2 height 0.1), circular orbits. */
static void make_disk(int n, float4* pos, float3* vel) {
    const double totalMass = 1.0;
    for (int i = 0; i < n; ++i) {
        double R;
        do {
            R = -log(uniform01() * uniform01() + 1e-300);
        } while (R > 5.0);
        const double phi = 2.0 * M_PI * uniform01();
        double u = uniform01();
        if (u < 1e-12) u = 1e-12;
        if (u > 1 - 1e-12) u = 1 - 1e-12;
        const double z = 0.1 * atanh(2.0 * u - 1.0);
        /* enclosed mass of a truncated exponential disk */
        const double menc = totalMass * (1.0 - (1.0 + R) * exp(-R)) / (1.0 - 6.0 * exp(-5.0));
        const double vc = sqrt(menc / (R > 1e-3 ? R : 1e-3));
        pos[i].x = (float)(R * cos(phi));
        pos[i].y = (float)(R * sin(phi));
        pos[i].z = (float)z;
        pos[i].w = (float)(totalMass / n);
        vel[i].x = (float)(-vc * sin(phi));
        vel[i].y = (float)(vc * cos(phi));
        vel[i].z = 0.0f;
    }
}

/* Reads an initial-conditions file into freshly allocated, zero-padded arrays. */
static int read_ic_file(const char* name, int* n, float4** pos, float3** vel, int padTo) {
    FILE* f = fopen(name, "rb");
    if (!f) return -1;
    int count = 0;
    int ok = fread(&count, sizeof(int), 1, f) == 1 && count > 0;
    if (ok) {
        const int padded = (count + padTo - 1) / padTo * padTo;
        *pos = (float4*)calloc((size_t)padded, sizeof(float4));
        *vel = (float3*)calloc((size_t)padded, sizeof(float3));
        ok = *pos && *vel && fread(*pos, sizeof(float4), (size_t)count, f) == (size_t)count &&
             fread(*vel, sizeof(float3), (size_t)count, f) == (size_t)count;
    }
    fclose(f);
    *n = count;
    return ok ? 0 : -1;
}

int main(int argc, char** argv) {
    int starCount = argc > 1 ? atoi(argv[1]) : 16384;
    const int tSteps = argc > 2 ? atoi(argv[2]) : 100;
    const float th = argc > 3 ? (float)atof(argv[3]) : theta;
    const float eps = argc > 4 ? (float)atof(argv[4]) : 0.01f;
    const float dt = argc > 5 ? (float)atof(argv[5]) : 1e-3f;
    const int framesPerWrite = argc > 6 ? atoi(argv[6]) : 5;
    const char* outName = argc > 7 ? argv[7] : "outputDump.bin";

    float4* hostPos = NULL;
    float3* hostVel = NULL;
    if (argc > 8 && read_ic_file(argv[8], &starCount, &hostPos, &hostVel, WARPTHREADS) != 0) {
        printf("Could not read initial conditions from %s\n", argv[8]);
        return -1;
    }
    if (starCount <= 0) {
        printf("No bodies\n");
        return -1;
    }

    int accelerationBlocks, treeBuilderBlocks, integrateBlocks, integrationThreadsPerBlock,
        treeBuilderThreadsPerBlock, paddedStarCount;
    blockNumCalculator(&accelerationBlocks, &treeBuilderBlocks, &integrateBlocks,
                       &paddedStarCount, &integrationThreadsPerBlock,
                       &treeBuilderThreadsPerBlock, starCount);

    if (!hostPos) {
        hostPos = (float4*)calloc((size_t)paddedStarCount, sizeof(float4));
        hostVel = (float3*)calloc((size_t)paddedStarCount, sizeof(float3));
        make_disk(starCount, hostPos, hostVel);
    }
    uint32_t* ident = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)starCount);
    uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)starCount);
    uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)starCount);
    float* frame = (float*)malloc(sizeof(float) * 3 * (size_t)starCount);
    for (int i = 0; i < starCount; ++i) ident[i] = (uint32_t)i;

    float4* pos = (float4*)gd_malloc(sizeof(float4) * (size_t)paddedStarCount);
    float3* vel = (float3*)gd_malloc(sizeof(float3) * (size_t)paddedStarCount);
    float3* acc = (float3*)gd_malloc(sizeof(float3) * (size_t)paddedStarCount);
    gd_memcpy_h2d(pos, hostPos, sizeof(float4) * (size_t)paddedStarCount);
    gd_memcpy_h2d(vel, hostVel, sizeof(float3) * (size_t)paddedStarCount);
    gd_memset(acc, 0, sizeof(float3) * (size_t)paddedStarCount);

    treeState* state = (treeState*)calloc(1, sizeof(treeState));
    initialiseTree(state, starCount, paddedStarCount, integrateBlocks);

    FILE* out = fopen(outName, "wb");
    if (!out) {
        printf("File pointer null\n");
        return -1;
    }
    const int writeSteps = tSteps / framesPerWrite;
    fwrite(&starCount, sizeof(int), 1, out);
    fwrite(&writeSteps, sizeof(int), 1, out);
    RGB* colours = (RGB*)malloc(sizeof(RGB) * (size_t)starCount);
    for (int i = 0; i < starCount; ++i) colours[i].R = colours[i].G = colours[i].B = 100;
    fwrite(colours, sizeof(RGB), (size_t)starCount, out);

    void* e0 = gd_event_create();
    void* e1 = gd_event_create();
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int frames = 0;
    for (int step = 0; step < tSteps; ++step) {
        gd_event_record(e0);
        buildTree(state, &pos, &vel);
        calculateAcceleration(pos, *state, acc, eps * eps, G, th, starCount);
        positionUpdater(pos, vel, acc, starCount, dt);
        gd_event_record(e1);
        const float ms = gd_event_elapsed_ms(e0, e1);
        /* carry original identities through this step's sort */
        gd_memcpy_d2h(perm, state->indexOut, sizeof(uint32_t) * (size_t)starCount);
        for (int i = 0; i < starCount; ++i) tmp[i] = ident[perm[i]];
        memcpy(ident, tmp, sizeof(uint32_t) * (size_t)starCount);
        if (step % framesPerWrite == 0 && frames < writeSteps) {
            gd_memcpy_d2h(hostPos, pos, sizeof(float4) * (size_t)starCount);
            for (int i = 0; i < starCount; ++i) {
                frame[3 * ident[i]] = hostPos[i].x;
                frame[3 * ident[i] + 1] = hostPos[i].y;
                frame[3 * ident[i] + 2] = hostPos[i].z;
            }
            fwrite(frame, sizeof(float), 3 * (size_t)starCount, out);
            ++frames;
        }
        if (step == 0 || step == tSteps - 1)
            printf("step %d: %.3f ms, deepest level %d\n", step, ms, state->deepestLevel);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    fclose(out);
    const double secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    printf("bodies %d steps %d: %.3f s, %.3e body-steps/s, frames %d\n", starCount, tSteps, secs,
           (double)starCount * tSteps / secs, frames);
    return 0;
}
