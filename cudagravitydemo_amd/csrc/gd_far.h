/*
 * Far cells of a group's list through one local expansion per group.
 *
 * A group (32-ish Morton-consecutive bodies) sees hundreds of quadrupole
 * cells; for a cell whose centre of mass is far from the group's box
 * compared to the box's size (h / R < far, h the box's half diagonal, R the
 * distance from the box centre g to the cell), the field of the cell varies
 * slowly across the group.  Such cells are summed once, by one lane each,
 * into the coefficients of a Taylor expansion of the field about g:
 *
 *   a(g + d) = C1 + C2 d + C3 d d / 2 + C4 d d d / 6,
 *
 * C_n = sum over cells of m D_n(y) + S_ab D_(n+2)(y)_(..ab) / 2 for n <= 3 and
 * C4 = sum of m D_4(y), with y = g - c, m the cell's mass, S its second
 * moments about c (not trace-free: the softened kernel is not harmonic) and
 * D_n the n-th derivative tensor of the softened kernel K = (|y|^2 + eps^2)^-1/2.
 * Every body of the group then evaluates the expansion once.  Truncation:
 * O((h/R)^4) of the monopole and O((h/R)^3) of the quadrupole terms.
 *
 * Symmetric tensors are stored by their distinct components:
 *   C1: x y z;  C2: xx yy zz xy xz yz;
 *   C3: xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz;
 *   C4: xxxx yyyy zzzz xxxy xxxz xyyy yyyz xzzz yzzz xxyy xxzz yyzz xxyz xyyz xyzz.
 */
#ifndef GD_FAR_H
#define GD_FAR_H

#define GD_LOCAL_TERMS 35   /* 34 coefficients + the sum of the cells' 3 m r^-5 */

struct gd_local {
    float c[GD_LOCAL_TERMS];
};

__device__ __forceinline__ void gd_local_zero(gd_local& L) {
#pragma unroll
    for (int k = 0; k < GD_LOCAL_TERMS; ++k) L.c[k] = 0.f;
}

/* Adds the terms of one cell (mass m, second moments q0 = (Sxx Syy Szz Sxy),
 * q1 = (Sxz Syz . .)) with (x, y, z) = g - centre of mass.  Every term is an
 * FMA chain into its accumulator; the constant delta-delta part of C4 (3 m
 * r^-5 per cell) is summed once, in c[34], and spread by gd_local_finish. */
__device__ __forceinline__ void gd_local_add(gd_local& L, float x, float y, float z, float m,
                                             float4 q0, float4 q1, float eps2) {
    float* c = L.c;
    const float Sxx = q0.x, Syy = q0.y, Szz = q0.z, Sxy = q0.w, Sxz = q1.x, Syz = q1.y;
    const float r2 = __builtin_fmaf(x, x, __builtin_fmaf(y, y, __builtin_fmaf(z, z, eps2)));
    const float r1 = __builtin_amdgcn_rsqf(r2);
    const float w = r1 * r1;
    const float r3 = r1 * w, r5 = r3 * w, r7 = r5 * w, r9 = r7 * w, r11 = r9 * w;
    const float sx = __builtin_fmaf(Sxx, x, __builtin_fmaf(Sxy, y, Sxz * z));
    const float sy = __builtin_fmaf(Sxy, x, __builtin_fmaf(Syy, y, Syz * z));
    const float sz = __builtin_fmaf(Sxz, x, __builtin_fmaf(Syz, y, Szz * z));
    const float q = __builtin_fmaf(x, sx, __builtin_fmaf(y, sy, z * sz));
    const float s = Sxx + Syy + Szz;
    const float mr3 = m * r3, mr5 = m * r5, mr7 = m * r7, mr9 = m * r9;
    /* coefficient of y (C1) and of the Kronecker delta (C2) */
    const float a1 = __builtin_fmaf(1.5f * s, r5, -__builtin_fmaf(7.5f * q, r7, mr3));
    /* of S y (C1) and of S (C2) */
    const float f5 = 3.f * r5;
    /* of y y (C2) and of the symmetrised delta y (C3) */
    const float b2 = __builtin_fmaf(3.f, mr5, __builtin_fmaf(52.5f * q, r9, -7.5f * (s * r7)));
    /* of S y y + y S y (C2), the symmetrised delta S y and y S (C3) */
    const float e7 = -15.f * r7;
    /* of y y y (C3) and of the symmetrised S y y (C3) */
    const float al3 = __builtin_fmaf(52.5f * s, r9, -__builtin_fmaf(472.5f * q, r11, 15.f * mr7));
    const float g9 = 105.f * r9;
    /* monopole fourth-rank terms: y^4 and the symmetrised delta y y */
    const float A4 = 105.f * mr9, B4 = -15.f * mr7;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
    const float u = e7 * sx, v = e7 * sy, t = e7 * sz;
    const float x2 = x + x, y2 = y + y, z2 = z + z;
    /* C1 = a1 y + f5 S y */
    c[0] = __builtin_fmaf(a1, x, __builtin_fmaf(f5, sx, c[0]));
    c[1] = __builtin_fmaf(a1, y, __builtin_fmaf(f5, sy, c[1]));
    c[2] = __builtin_fmaf(a1, z, __builtin_fmaf(f5, sz, c[2]));
    /* C2 = b2 y y + a1 delta + e7 (S y y + y S y) + f5 S */
    c[3] = __builtin_fmaf(b2, xx, __builtin_fmaf(u, x2, __builtin_fmaf(f5, Sxx, c[3] + a1)));
    c[4] = __builtin_fmaf(b2, yy, __builtin_fmaf(v, y2, __builtin_fmaf(f5, Syy, c[4] + a1)));
    c[5] = __builtin_fmaf(b2, zz, __builtin_fmaf(t, z2, __builtin_fmaf(f5, Szz, c[5] + a1)));
    c[6] = __builtin_fmaf(b2, xy, __builtin_fmaf(u, y, __builtin_fmaf(v, x, __builtin_fmaf(f5, Sxy, c[6]))));
    c[7] = __builtin_fmaf(b2, xz, __builtin_fmaf(u, z, __builtin_fmaf(t, x, __builtin_fmaf(f5, Sxz, c[7]))));
    c[8] = __builtin_fmaf(b2, yz, __builtin_fmaf(v, z, __builtin_fmaf(t, y, __builtin_fmaf(f5, Syz, c[8]))));
    /* C3 = al3 y y y + sym(delta p) + g9 sym(S y y y) + e7 sym(y S), with
       p = b2 y + e7 S y; sym(.) sums the distinct index placements */
    const float px = __builtin_fmaf(b2, x, u), py = __builtin_fmaf(b2, y, v),
                pz = __builtin_fmaf(b2, z, t);
    const float gx = g9 * sx, gy = g9 * sy, gz = g9 * sz;
    const float ax = al3 * x, ay = al3 * y, az = al3 * z;
    const float ex = e7 * x, ey = e7 * y, ez = e7 * z;
    const float Sxy2 = Sxy + Sxy, Sxz2 = Sxz + Sxz, Syz2 = Syz + Syz;
    const float xy2 = xy + xy, xz2 = xz + xz, yz2 = yz + yz;
    c[9] = __builtin_fmaf(3.f, __builtin_fmaf(gx, xx, __builtin_fmaf(ex, Sxx, px)),
                          __builtin_fmaf(ax, xx, c[9]));
    c[10] = __builtin_fmaf(3.f, __builtin_fmaf(gy, yy, __builtin_fmaf(ey, Syy, py)),
                           __builtin_fmaf(ay, yy, c[10]));
    c[11] = __builtin_fmaf(3.f, __builtin_fmaf(gz, zz, __builtin_fmaf(ez, Szz, pz)),
                           __builtin_fmaf(az, zz, c[11]));
    /* xxy */
    c[12] = __builtin_fmaf(ax, xy, __builtin_fmaf(gx, xy2, __builtin_fmaf(gy, xx,
            __builtin_fmaf(ex, Sxy2, __builtin_fmaf(ey, Sxx, c[12] + py)))));
    /* xxz */
    c[13] = __builtin_fmaf(ax, xz, __builtin_fmaf(gx, xz2, __builtin_fmaf(gz, xx,
            __builtin_fmaf(ex, Sxz2, __builtin_fmaf(ez, Sxx, c[13] + pz)))));
    /* xyy */
    c[14] = __builtin_fmaf(ay, xy, __builtin_fmaf(gy, xy2, __builtin_fmaf(gx, yy,
            __builtin_fmaf(ey, Sxy2, __builtin_fmaf(ex, Syy, c[14] + px)))));
    /* yyz */
    c[15] = __builtin_fmaf(ay, yz, __builtin_fmaf(gy, yz2, __builtin_fmaf(gz, yy,
            __builtin_fmaf(ey, Syz2, __builtin_fmaf(ez, Syy, c[15] + pz)))));
    /* xzz */
    c[16] = __builtin_fmaf(az, xz, __builtin_fmaf(gz, xz2, __builtin_fmaf(gx, zz,
            __builtin_fmaf(ez, Sxz2, __builtin_fmaf(ex, Szz, c[16] + px)))));
    /* yzz */
    c[17] = __builtin_fmaf(az, yz, __builtin_fmaf(gz, yz2, __builtin_fmaf(gy, zz,
            __builtin_fmaf(ez, Syz2, __builtin_fmaf(ey, Szz, c[17] + py)))));
    /* xyz */
    c[18] = __builtin_fmaf(ax, yz, __builtin_fmaf(gx, yz, __builtin_fmaf(gy, xz,
            __builtin_fmaf(gz, xy, __builtin_fmaf(ex, Syz, __builtin_fmaf(ey, Sxz,
            __builtin_fmaf(ez, Sxy, c[18])))))));
    /* C4 = A4 y y y y + B4 sym(delta y y) + 3 m r^-5 sym(delta delta) */
    const float Axx = A4 * xx, Ayy = A4 * yy, Azz = A4 * zz;
    const float B3 = 3.f * B4, B6 = 6.f * B4;
    const float ax3 = Axx + B3, ay3 = Ayy + B3, az3 = Azz + B3;
    const float axb = Axx + B4, ayb = Ayy + B4, azb = Azz + B4;
    c[19] = __builtin_fmaf(Axx, xx, __builtin_fmaf(B6, xx, c[19]));
    c[20] = __builtin_fmaf(Ayy, yy, __builtin_fmaf(B6, yy, c[20]));
    c[21] = __builtin_fmaf(Azz, zz, __builtin_fmaf(B6, zz, c[21]));
    c[22] = __builtin_fmaf(ax3, xy, c[22]);
    c[23] = __builtin_fmaf(ax3, xz, c[23]);
    c[24] = __builtin_fmaf(ay3, xy, c[24]);
    c[25] = __builtin_fmaf(ay3, yz, c[25]);
    c[26] = __builtin_fmaf(az3, xz, c[26]);
    c[27] = __builtin_fmaf(az3, yz, c[27]);
    c[28] = __builtin_fmaf(axb, yy, __builtin_fmaf(B4, xx, c[28]));
    c[29] = __builtin_fmaf(axb, zz, __builtin_fmaf(B4, xx, c[29]));
    c[30] = __builtin_fmaf(ayb, zz, __builtin_fmaf(B4, yy, c[30]));
    c[31] = __builtin_fmaf(axb, yz, c[31]);
    c[32] = __builtin_fmaf(ayb, xz, c[32]);
    c[33] = __builtin_fmaf(azb, xy, c[33]);
    c[34] = __builtin_fmaf(3.f, mr5, c[34]);
}

/* After the sum over cells: the delta-delta part of C4 (sum of 3 m r^-5)
 * into its components, 3x for xxxx yyyy zzzz, 1x for xxyy xxzz yyzz. */
__device__ __forceinline__ void gd_local_finish(gd_local& L) {
    float* c = L.c;
    const float d = c[34], d3 = 3.f * d;
    c[19] += d3;
    c[20] += d3;
    c[21] += d3;
    c[28] += d;
    c[29] += d;
    c[30] += d;
}

#define GD_LOCAL_STRIDE 36   /* floats per lane in the reduction buffer (16-byte rows) */

/* Sum of every lane's coefficients, the same in every lane, through LDS
 * (sl: 64 x GD_LOCAL_STRIDE floats): each lane writes its row, lane k sums
 * column k over the 64 rows in lane order (a fixed order: the result does not
 * depend on timing), and every lane reads the sums back. */
__device__ __forceinline__ void gd_local_allreduce(gd_local& L, float* sl) {
    const unsigned lane = gd_lane();
    float4* row = reinterpret_cast<float4*>(sl + lane * GD_LOCAL_STRIDE);
#pragma unroll
    for (int k = 0; k < GD_LOCAL_STRIDE / 4; ++k) {
        const int b = 4 * k;
        row[k] = make_float4(L.c[b], b + 1 < GD_LOCAL_TERMS ? L.c[b + 1] : 0.f,
                             b + 2 < GD_LOCAL_TERMS ? L.c[b + 2] : 0.f,
                             b + 3 < GD_LOCAL_TERMS ? L.c[b + 3] : 0.f);
    }
    __builtin_amdgcn_wave_barrier();
    float sum = 0.f;
    if (lane < GD_LOCAL_TERMS) {
#pragma unroll 16
        for (int r = 0; r < 64; ++r) sum += sl[r * GD_LOCAL_STRIDE + lane];
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < GD_LOCAL_TERMS) sl[lane] = sum;
    __builtin_amdgcn_wave_barrier();
    const float4* tot = reinterpret_cast<const float4*>(sl);
#pragma unroll
    for (int k = 0; k < GD_LOCAL_STRIDE / 4; ++k) {
        const float4 t = tot[k];
        const int b = 4 * k;
        L.c[b] = t.x;
        if (b + 1 < GD_LOCAL_TERMS) L.c[b + 1] = t.y;
        if (b + 2 < GD_LOCAL_TERMS) L.c[b + 2] = t.z;
        if (b + 3 < GD_LOCAL_TERMS) L.c[b + 3] = t.w;
    }
    __builtin_amdgcn_wave_barrier();
}

/* The expansion at offset (dx, dy, dz) from its centre, added to (ax, ay, az). */
__device__ __forceinline__ void gd_local_eval(const gd_local& L, float dx, float dy, float dz,
                                              float& ax, float& ay, float& az) {
    const float* c = L.c;
    /* second order: C3 / 2 against the monomials, the mixed ones doubled */
    const float xx = 0.5f * dx * dx, yy = 0.5f * dy * dy, zz = 0.5f * dz * dz;
    const float xy = dx * dy, xz = dx * dz, yz = dy * dz;
    /* third order: C4 / 6, multiplicities 1, 3, 6 */
    const float s6 = 1.f / 6.f;
    const float x3 = s6 * dx * dx * dx, y3 = s6 * dy * dy * dy, z3 = s6 * dz * dz * dz;
    const float x2y = 0.5f * dx * xy, x2z = 0.5f * dx * xz, xy2 = 0.5f * dy * xy;
    const float y2z = 0.5f * dy * yz, xz2 = 0.5f * dz * xz, yz2 = 0.5f * dz * yz;
    const float xyz = xy * dz;
    float x = c[0], y = c[1], z = c[2];
    x = __builtin_fmaf(c[3], dx, __builtin_fmaf(c[6], dy, __builtin_fmaf(c[7], dz, x)));
    y = __builtin_fmaf(c[6], dx, __builtin_fmaf(c[4], dy, __builtin_fmaf(c[8], dz, y)));
    z = __builtin_fmaf(c[7], dx, __builtin_fmaf(c[8], dy, __builtin_fmaf(c[5], dz, z)));
    /* C3: xxx 9, yyy 10, zzz 11, xxy 12, xxz 13, xyy 14, yyz 15, xzz 16, yzz 17, xyz 18 */
    x = __builtin_fmaf(c[9], xx, __builtin_fmaf(c[14], yy, __builtin_fmaf(c[16], zz, x)));
    x = __builtin_fmaf(c[12], xy, __builtin_fmaf(c[13], xz, __builtin_fmaf(c[18], yz, x)));
    y = __builtin_fmaf(c[12], xx, __builtin_fmaf(c[10], yy, __builtin_fmaf(c[17], zz, y)));
    y = __builtin_fmaf(c[14], xy, __builtin_fmaf(c[18], xz, __builtin_fmaf(c[15], yz, y)));
    z = __builtin_fmaf(c[13], xx, __builtin_fmaf(c[15], yy, __builtin_fmaf(c[11], zz, z)));
    z = __builtin_fmaf(c[18], xy, __builtin_fmaf(c[16], xz, __builtin_fmaf(c[17], yz, z)));
    /* C4: xxxx 19, yyyy 20, zzzz 21, xxxy 22, xxxz 23, xyyy 24, yyyz 25, xzzz 26,
       yzzz 27, xxyy 28, xxzz 29, yyzz 30, xxyz 31, xyyz 32, xyzz 33 */
    x = __builtin_fmaf(c[19], x3, __builtin_fmaf(c[24], y3, __builtin_fmaf(c[26], z3, x)));
    x = __builtin_fmaf(c[22], x2y, __builtin_fmaf(c[23], x2z, __builtin_fmaf(c[28], xy2, x)));
    x = __builtin_fmaf(c[29], xz2, __builtin_fmaf(c[32], y2z, __builtin_fmaf(c[33], yz2, x)));
    x = __builtin_fmaf(c[31], xyz, x);
    y = __builtin_fmaf(c[22], x3, __builtin_fmaf(c[20], y3, __builtin_fmaf(c[27], z3, y)));
    y = __builtin_fmaf(c[28], x2y, __builtin_fmaf(c[31], x2z, __builtin_fmaf(c[24], xy2, y)));
    y = __builtin_fmaf(c[33], xz2, __builtin_fmaf(c[25], y2z, __builtin_fmaf(c[30], yz2, y)));
    y = __builtin_fmaf(c[32], xyz, y);
    z = __builtin_fmaf(c[23], x3, __builtin_fmaf(c[25], y3, __builtin_fmaf(c[21], z3, z)));
    z = __builtin_fmaf(c[31], x2y, __builtin_fmaf(c[29], x2z, __builtin_fmaf(c[32], xy2, z)));
    z = __builtin_fmaf(c[26], xz2, __builtin_fmaf(c[30], y2z, __builtin_fmaf(c[27], yz2, z)));
    z = __builtin_fmaf(c[33], xyz, z);
    ax += x;
    ay += y;
    az += z;
}

#endif
