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
