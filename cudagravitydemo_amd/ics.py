"""Seeded initial conditions for the benchmark configurations.

Generated on the host in float64 with numpy's PCG64 and rounded to float32,
so the CPU oracle and the GPU engine receive identical bodies.  Layout:
pos_mass (N, 4) float32 = x, y, z, m; vel (N, 3) float32.

Choices the configurations leave open (recorded here and in the design notes):
  * the central point mass of a disk is body 0 of that disk (mass 0.1); the
    other N-1 bodies share the disk mass 1 equally, so N stays a power of two;
  * radial positions follow Sigma ~ exp(-R/h): R = -h ln(u1 u2), redrawn
    above 5 h (the reference's truncation, initialConditions.c:20);
  * vertical positions follow sech^2(z/z0): z = z0 atanh(2u - 1);
  * circular speed from the enclosed mass of the sampled bodies (spherical
    radius, central mass included), v_c^2 = G M(<r) / r; isotropic Gaussian
    dispersion of 5 % of v_c per component;
  * the merger places two such disks 20 h apart on a parabolic orbit with
    pericentre 2 h, the second disk inclined by 30 degrees about the x axis;
  * the Plummer sphere uses Aarseth, Henon & Wielen (1974): r = a (u^{-2/3}
    - 1)^{-1/2} (redrawn beyond 20 a), speeds by rejection from
    q^2 (1 - q^2)^{7/2}, isotropic directions.
"""
import numpy as np

G_UNIT = 1.0


def _disk(rng, n, h=1.0, z0=0.1, disk_mass=1.0, central_mass=0.1, dispersion=0.05, G=G_UNIT):
    """One exponential disk with a central point mass as body 0 (float64)."""
    nd = n - 1
    R = np.empty(nd)
    todo = np.arange(nd)
    while todo.size:
        r = -h * np.log(rng.random(todo.size) * rng.random(todo.size) + 1e-300)
        ok = r <= 5.0 * h
        R[todo[ok]] = r[ok]
        todo = todo[~ok]
    phi = rng.random(nd) * 2.0 * np.pi
    u = rng.random(nd)
    u = np.clip(u, 1e-12, 1 - 1e-12)
    z = z0 * np.arctanh(2.0 * u - 1.0)
    pos = np.zeros((n, 3))
    pos[1:, 0] = R * np.cos(phi)
    pos[1:, 1] = R * np.sin(phi)
    pos[1:, 2] = z
    m = np.empty(n)
    m[0] = central_mass
    m[1:] = disk_mass / nd
    # enclosed mass by spherical radius
    rs = np.sqrt(R * R + z * z)
    order = np.argsort(rs, kind="stable")
    cum = np.empty(nd)
    cum[order] = np.cumsum(m[1:][order]) - m[1:][order] * 0.5
    menc = central_mass + cum
    vc = np.sqrt(G * menc / np.maximum(rs, 1e-6))
    vel = np.zeros((n, 3))
    vel[1:, 0] = -vc * np.sin(phi)
    vel[1:, 1] = vc * np.cos(phi)
    vel[1:] += rng.standard_normal((nd, 3)) * (dispersion * vc)[:, None]
    return pos, vel, m


def _pack(pos, vel, m):
    pm = np.empty((pos.shape[0], 4), dtype=np.float32)
    pm[:, :3] = pos
    pm[:, 3] = m
    return pm, np.ascontiguousarray(vel, dtype=np.float32)


def exponential_disk(n=16384, seed=1234, h=1.0, z0=0.1, central_mass=0.1, dispersion=0.05):
    """Config 0: one disk, N bodies (body 0 = central mass 0.1), total disk mass 1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos, vel, m = _disk(rng, n, h=h, z0=z0, central_mass=central_mass, dispersion=dispersion)
    return _pack(pos, vel, m)


def _rot_x(deg):
    a = np.deg2rad(deg)
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def disk_merger(n_per_disk=1 << 20, seed=2024, separation=20.0, pericentre=2.0, inclination=30.0,
                h=1.0, z0=0.1, central_mass=0.1, dispersion=0.05, G=G_UNIT):
    """Config 1: two disks, inclined, on a parabolic approach orbit."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pa, va, ma = _disk(rng, n_per_disk, h=h, z0=z0, central_mass=central_mass,
                       dispersion=dispersion, G=G)
    pb, vb, mb = _disk(rng, n_per_disk, h=h, z0=z0, central_mass=central_mass,
                       dispersion=dispersion, G=G)
    rot = _rot_x(inclination)
    pb = pb @ rot.T
    vb = vb @ rot.T
    mtot = ma.sum() + mb.sum()
    d = separation
    v2 = 2.0 * G * mtot / d                    # parabolic: E = 0
    vt = np.sqrt(2.0 * G * mtot * pericentre) / d
    vr = np.sqrt(max(v2 - vt * vt, 0.0))
    rel_pos = np.array([d, 0.0, 0.0])          # B relative to A
    rel_vel = np.array([-vr, vt, 0.0])         # approaching, with angular momentum
    fa = mb.sum() / mtot
    fb = ma.sum() / mtot
    pa = pa - fa * rel_pos
    va = va - fa * rel_vel
    pb = pb + fb * rel_pos
    vb = vb + fb * rel_vel
    pos = np.concatenate([pa, pb])
    vel = np.concatenate([va, vb])
    m = np.concatenate([ma, mb])
    return _pack(pos, vel, m)


def plummer(n=1 << 22, seed=7, a=1.0, M=1.0, G=G_UNIT, rmax=20.0):
    """Config 2: Plummer sphere (Aarseth, Henon & Wielen 1974), equal masses."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        u = rng.random(todo.size)
        rr = a / np.sqrt(np.maximum(u, 1e-300) ** (-2.0 / 3.0) - 1.0)
        ok = rr <= rmax * a
        r[todo[ok]] = rr[ok]
        todo = todo[~ok]

    def iso(k):
        cth = 2.0 * rng.random(k) - 1.0
        ph = 2.0 * np.pi * rng.random(k)
        sth = np.sqrt(1.0 - cth * cth)
        return np.stack([sth * np.cos(ph), sth * np.sin(ph), cth], axis=1)

    pos = iso(n) * r[:, None]
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x = rng.random(todo.size)
        y = rng.random(todo.size) * 0.1
        ok = y < x * x * (1.0 - x * x) ** 3.5
        q[todo[ok]] = x[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0 * G * M / a) * (1.0 + (r / a) ** 2) ** -0.25
    vel = iso(n) * (q * vesc)[:, None]
    m = np.full(n, M / n)
    # centre of mass frame
    pos -= (m[:, None] * pos).sum(0) / M
    vel -= (m[:, None] * vel).sum(0) / M
    return _pack(pos, vel, m)


def uniform_cube(n, seed=0, size=1.0):
    """Small test input: uniform cube, equal masses summing to 1, zero velocities."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.random((n, 3)) * size
    return _pack(pos, np.zeros((n, 3)), np.full(n, 1.0 / n))


def subsample(n, k=4096, seed=99):
    """Seeded subsample of k body indices (sorted, without replacement)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = min(k, n)
    return np.sort(rng.choice(n, size=k, replace=False)).astype(np.int32)


def rotation_matrix(yaw, pitch, roll):
    """The reference's disk rotation (initialConditions.c:199-232).  Its three
    matrices are written as cglm literals, which cglm stores column by column,
    so each literal is the transpose of how it reads; the product
    yaw * pitch * roll (glm_mat3_mul) is then Rz(-yaw) Ry(-pitch) Rx(-roll)
    with Rz, Ry, Rx the usual right-handed rotations."""
    cy, sy = np.cos(yaw), np.sin(yaw)
    cp, sp = np.cos(pitch), np.sin(pitch)
    cr, sr = np.cos(roll), np.sin(roll)
    rz = np.array([[cy, sy, 0.0], [-sy, cy, 0.0], [0.0, 0.0, 1.0]])
    ry = np.array([[cp, 0.0, -sp], [0.0, 1.0, 0.0], [sp, 0.0, cp]])
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cr, sr], [0.0, -sr, cr]])
    return rz @ ry @ rx


def reference_scenario(n=100000, seed=0, scale_length=4.0, scale_thickness=0.3, total_mass=5.0,
                       halo_fraction=0.2, angles=(0.0, 0.0, 0.0), centre=(0.0, 0.0, 0.0),
                       velocity=(0.0, 0.0, 0.0), G=G_UNIT):
    """The reference's own run (basicGravity.cu:517-545, 694-697): N bodies of
    equal mass total_mass / N; the first N - int(0.2 N) form the disk of
    generateLocalCoordsDisk (initialConditions.c:246-302), the rest the
    'halo' of generateDarkMatterHalo (initialConditions.c:314-382).

    Disk: radius by rejection sampling of R exp(-R/h) under its maximum h/e
    on [0, 5h] (sampleRadius, :44-69); uniform azimuth; height from the
    Laplace profile exp(-|z|/z0) by inversion (sampleHeight, :150-162);
    velocity the circular speed sqrt(G M(<R) / R) of an exponential disk of
    total_mass, M(<R) = M (1 - (1 + R/h) exp(-R/h)) (:164-176), tangential,
    no vertical motion; then rotated by rotation_matrix(*angles) and offset by
    centre / velocity.  Halo: radius uniform in [0, 5h], direction uniform on
    the sphere (cos theta uniform), at rest relative to `velocity` (:314-361).
    The reference draws from rand() seeded with the time; this draws from a
    seeded PCG64 instead."""
    rng = np.random.Generator(np.random.PCG64(seed))
    halo = int(n * halo_fraction)
    nd = n - halo
    h = scale_length
    R = np.empty(nd)
    todo = np.arange(nd)
    ymax = h * np.exp(-1.0)
    while todo.size:
        r = 5.0 * h * rng.random(todo.size)
        y = ymax * rng.random(todo.size)
        ok = y < r * np.exp(-r / h)
        R[todo[ok]] = r[ok]
        todo = todo[~ok]
    phi = 2.0 * np.pi * rng.random(nd)
    u = rng.random(nd)
    u = np.clip(u, 1e-300, 1.0 - 1e-16)
    z = np.where(u >= 0.5, -np.log(2.0 - 2.0 * u), np.log(2.0 * u)) * scale_thickness
    menc = total_mass * (1.0 - (1.0 + R / h) * np.exp(-R / h))
    vc = np.sqrt(G * menc / np.maximum(R, 1e-30))
    pd = np.stack([R * np.cos(phi), R * np.sin(phi), z], 1)
    vd = np.stack([-vc * np.sin(phi), vc * np.cos(phi), np.zeros(nd)], 1)
    rot = rotation_matrix(*angles)
    pd = pd @ rot.T + np.asarray(centre, float)
    vd = vd @ rot.T + np.asarray(velocity, float)
    cth = 2.0 * rng.random(halo) - 1.0
    ph = 2.0 * np.pi * rng.random(halo)
    rh = 5.0 * h * rng.random(halo)
    sth = np.sqrt(1.0 - cth * cth)
    phalo = np.stack([rh * sth * np.cos(ph), rh * sth * np.sin(ph), rh * cth], 1) + \
        np.asarray(centre, float)
    vhalo = np.tile(np.asarray(velocity, float), (halo, 1))
    pos = np.concatenate([pd, phalo])
    vel = np.concatenate([vd, vhalo])
    return _pack(pos, vel, np.full(n, total_mass / n))


def write_ic_file(path, pos_mass, vel):
    """Initial conditions for the C driver: int32 N, float32 (x, y, z, m) x N,
    float32 (vx, vy, vz) x N, little-endian."""
    pos_mass = np.ascontiguousarray(pos_mass, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    with open(path, "wb") as f:
        f.write(np.int32(pos_mass.shape[0]).tobytes())
        f.write(pos_mass.tobytes())
        f.write(vel.tobytes())


def read_trajectory(path):
    """The reference's output file (basicGravity.cu:614-635, 793-801): int32 N,
    int32 frames, RGB[N] (3 bytes each), then float32[3 N] per frame.
    Returns (N, colours (N, 3) uint8, frames (F, N, 3) float32)."""
    raw = np.fromfile(path, dtype=np.uint8)
    n, frames = np.frombuffer(raw[:8].tobytes(), np.int32)
    colours = raw[8:8 + 3 * n].reshape(n, 3)
    body = raw[8 + 3 * n:]
    assert body.size == 12 * n * frames, (body.size, n, frames)
    return int(n), colours, np.frombuffer(body.tobytes(), np.float32).reshape(frames, n, 3)
