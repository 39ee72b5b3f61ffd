/*
 * gd_integrate.hip -- kick-drift, hi/lo positions, FP64 judge, FP64 diagnostics.
 *
 * positionUpdater restates tree.cu:678-703: v = fmaf(a, dt, v), then
 * x = fmaf(v_new, dt, x) (the oracle's operation order), one body per thread,
 * 40 B read and 24 B written per body.
 *
 * The hi/lo integrator keeps each coordinate as an unevaluated sum hi + lo
 * of two floats: the drift v*dt is added with an error-free two-sum and the
 * pair renormalised, so the low bits lost by a float add stay in lo
 * (Dekker 1971 / Knuth TAOCP 4.2.2 two-sum).
 *
 * The judge is an FP64 softened direct sum over all bodies for a sample of
 * bodies: one block per sample body, threads stride over sources, FP64
 * accumulation, wave then block reduction.
 */
#include "gd_internal.h"

#include <math.h>

__global__ __launch_bounds__(256) void k_integrate(float4* __restrict__ pos,
                                                   float3* __restrict__ vel,
                                                   const float3* __restrict__ acc, int n,
                                                   float dt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float3 a = acc[i];
    float3 v = vel[i];
    float4 p = pos[i];
    v.x = __builtin_fmaf(a.x, dt, v.x);
    v.y = __builtin_fmaf(a.y, dt, v.y);
    v.z = __builtin_fmaf(a.z, dt, v.z);
    vel[i] = v;
    p.x = __builtin_fmaf(v.x, dt, p.x);
    p.y = __builtin_fmaf(v.y, dt, p.y);
    p.z = __builtin_fmaf(v.z, dt, p.z);
    pos[i] = p;
}

extern "C" void positionUpdater(float4* positionMassArray, float3* velocityVals,
                                float3* accelerationVals, int starCount, float timeStep) {
    if (starCount <= 0) return;
    k_integrate<<<gd_div_up(starCount, 256), 256>>>(positionMassArray, velocityVals,
                                                    accelerationVals, starCount, timeStep);
    GD_LAUNCHED();
}

/* (hi, lo) += d, renormalised: error-free two-sum of hi and d, then the
 * rounding error and the old lo are folded back with a fast two-sum. */
__device__ __forceinline__ void gd_hilo_add(float& hi, float& lo, float d) {
    const float s = hi + d;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (d - bb);
    const float t = err + lo;
    const float h = s + t;
    lo = t - (h - s);
    hi = h;
}

__global__ __launch_bounds__(256) void k_integrate_hilo(float4* __restrict__ pos,
                                                        float3* __restrict__ lo,
                                                        float3* __restrict__ vel,
                                                        const float3* __restrict__ acc, int n,
                                                        float dt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float3 a = acc[i];
    float3 v = vel[i];
    float4 p = pos[i];
    float3 l = lo[i];
    v.x = __builtin_fmaf(a.x, dt, v.x);
    v.y = __builtin_fmaf(a.y, dt, v.y);
    v.z = __builtin_fmaf(a.z, dt, v.z);
    vel[i] = v;
    gd_hilo_add(p.x, l.x, v.x * dt);
    gd_hilo_add(p.y, l.y, v.y * dt);
    gd_hilo_add(p.z, l.z, v.z * dt);
    pos[i] = p;
    lo[i] = l;
}

extern "C" void gd_integrate_hilo(float4* pos, float3* posLo, float3* vel, const float3* acc,
                                  int starCount, float dt) {
    if (starCount <= 0) return;
    k_integrate_hilo<<<gd_div_up(starCount, 256), 256>>>(pos, posLo, vel, acc, starCount, dt);
    GD_LAUNCHED();
}

/* ---------------------------------------------------------------------------
 * FP64 direct-sum judge
 * ------------------------------------------------------------------------- */
__device__ __forceinline__ double gd_wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(256) void k_direct_fp64(const float4* __restrict__ pos,
                                                     const float3* __restrict__ lo, int n,
                                                     const int* __restrict__ sample,
                                                     double eps2, double G,
                                                     double* __restrict__ out) {
    __shared__ double sPart[3][4];
    const int s = sample[blockIdx.x];
    const float4 pi = pos[s];
    double xi = pi.x, yi = pi.y, zi = pi.z;
    if (lo) {
        const float3 li = lo[s];
        xi += li.x; yi += li.y; zi += li.z;
    }
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        if (j == s) continue;
        const float4 pj = pos[j];
        double xj = pj.x, yj = pj.y, zj = pj.z;
        if (lo) {
            const float3 lj = lo[j];
            xj += lj.x; yj += lj.y; zj += lj.z;
        }
        const double dx = xi - xj, dy = yi - yj, dz = zi - zj;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double inv = 1.0 / sqrt(r2);
        const double f = -G * (double)pj.w * inv * inv * inv;
        ax += f * dx;
        ay += f * dy;
        az += f * dz;
    }
    ax = gd_wave_sum_d(ax);
    ay = gd_wave_sum_d(ay);
    az = gd_wave_sum_d(az);
    const int wave = threadIdx.x >> 6;
    if (gd_lane() == 0) {
        sPart[0][wave] = ax;
        sPart[1][wave] = ay;
        sPart[2][wave] = az;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const double v = sPart[threadIdx.x][0] + sPart[threadIdx.x][1] + sPart[threadIdx.x][2] +
                         sPart[threadIdx.x][3];
        out[3 * blockIdx.x + threadIdx.x] = v;
    }
}

extern "C" void gd_direct_sum_fp64(const float4* pos, const float3* posLo, int starCount,
                                   const int* sample, int ns, double eps2, double G, double* acc) {
    if (ns <= 0 || starCount <= 0) return;
    k_direct_fp64<<<ns, 256>>>(pos, posLo, starCount, sample, eps2, G, acc);
    GD_LAUNCHED();
}

/* ---------------------------------------------------------------------------
 * FP64 diagnostics: kinetic energy, momentum, mass, potential energy
 * ------------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_diag(const float4* __restrict__ pos,
                                              const float3* __restrict__ vel,
                                              const float* __restrict__ phi, int n,
                                              double* __restrict__ out) {
    __shared__ double sPart[6][4];
    double ke = 0.0, px = 0.0, py = 0.0, pz = 0.0, m = 0.0, pe = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double mi = pos[i].w;
        const float3 v = vel[i];
        const double vx = v.x, vy = v.y, vz = v.z;
        ke += 0.5 * mi * (vx * vx + vy * vy + vz * vz);
        px += mi * vx;
        py += mi * vy;
        pz += mi * vz;
        m += mi;
        if (phi) pe += 0.5 * mi * (double)phi[i];
    }
    double v[6] = {ke, px, py, pz, m, pe};
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        v[k] = gd_wave_sum_d(v[k]);
        if (gd_lane() == 0) sPart[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const double s = sPart[threadIdx.x][0] + sPart[threadIdx.x][1] + sPart[threadIdx.x][2] +
                         sPart[threadIdx.x][3];
        atomicAdd(&out[threadIdx.x], s);
    }
}

extern "C" void gd_diagnostics(const float4* pos, const float3* vel, const float* phi,
                               int starCount, double* out) {
    double* d = nullptr;
    GD_HIP(hipMalloc(&d, 6 * sizeof(double)));
    if (gd_has_error()) return;
    GD_HIP(hipMemset(d, 0, 6 * sizeof(double)));
    if (starCount > 0) {
        int blocks = gd_div_up(starCount, 256);
        if (blocks > 1024) blocks = 1024;
        k_diag<<<blocks, 256>>>(pos, vel, phi, starCount, d);
        GD_LAUNCHED();
    }
    double h[6];
    GD_HIP(hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost));
    GD_HIP(hipFree(d));
    for (int k = 0; k < 6; ++k) out[k] = h[k];
    out[6] = h[0] + h[5];
}

/* ---------------------------------------------------------------------------
 * Stream copy of known size (16-byte vectors): a calibration kernel for
 * hardware byte counters and a bandwidth reference.
 * ------------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src,
                                                     float4* __restrict__ dst, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

extern "C" void gd_stream_copy(void* dst, const void* src, size_t bytes) {
    if (bytes % 16) {
        GD_FAIL("gd_stream_copy: %zu bytes is not a multiple of 16", bytes);
        return;
    }
    const long long n = (long long)(bytes / 16);
    if (n == 0) return;
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    k_stream_copy<<<(int)blocks, 256>>>((const float4*)src, (float4*)dst, n);
    GD_LAUNCHED();
}
