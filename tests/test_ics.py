"""The initial-condition generators against the distributions they sample
(CPU only): Kolmogorov-Smirnov tests against closed-form CDFs, exact
velocity relations, counts and masses, and the rotation against an
independent Euler-angle construction.

Reference scenario: initialConditions.c:44-69 (radius), :150-162 (height),
:164-176 (circular speed), :199-232 (rotation), :314-382 (halo), with the
run's parameters (basicGravity.cu:517-545, 694-697).
"""
import numpy as np
import pytest
from scipy import integrate, stats
from scipy.spatial.transform import Rotation

from cudagravitydemo_amd import ics

P_MIN = 1e-3   # a correct generator fails one of these KS tests with probability 1e-3


def ks(sample, cdf):
    return stats.kstest(np.asarray(sample, np.float64), cdf).pvalue


def test_rotation_matrix_is_the_cglm_product():
    # cglm literals are column-major: the product is Rz(-yaw) Ry(-pitch) Rx(-roll),
    # an intrinsic z-y-x rotation by the negated angles
    for yaw, pitch, roll in ((0.3, -0.7, 1.1), (0.0, 0.0, 0.0), (2.0, 0.4, -2.5)):
        want = Rotation.from_euler("ZYX", [-yaw, -pitch, -roll]).as_matrix()
        got = ics.rotation_matrix(yaw, pitch, roll)
        assert np.allclose(got, want, atol=1e-14)
        assert np.allclose(got @ got.T, np.eye(3), atol=1e-14)
    # a pure yaw turns x toward -y (the literal's first column is (cos, -sin, 0))
    r = ics.rotation_matrix(np.pi / 2, 0.0, 0.0)
    assert np.allclose(r @ [1.0, 0.0, 0.0], [0.0, -1.0, 0.0], atol=1e-15)


def test_reference_scenario_counts_and_masses():
    p, v = ics.reference_scenario(10001, seed=1)
    assert p.shape == (10001, 4) and v.shape == (10001, 3)
    assert np.all(p[:, 3] == np.float32(5.0 / 10001))
    halo = int(10001 * 0.2)
    assert np.all(v[10001 - halo:] == 0.0)          # the halo is at rest
    assert np.all(v[:10001 - halo, 2] == 0.0)       # no vertical motion in the disk plane


def test_reference_scenario_disk_distributions():
    n, h, z0, M = 200_000, 4.0, 0.3, 5.0
    p, v = ics.reference_scenario(n, seed=2)
    nd = n - int(0.2 * n)
    d = p[:nd, :3].astype(np.float64)
    R = np.hypot(d[:, 0], d[:, 1])
    norm = 1.0 - 6.0 * np.exp(-5.0)
    # R exp(-R/h) on [0, 5h]: CDF (1 - (1 + R/h) e^{-R/h}) / (1 - 6 e^{-5})
    assert ks(R, lambda r: np.clip((1 - (1 + r / h) * np.exp(-r / h)) / norm, 0, 1)) > P_MIN
    assert R.max() <= 5 * h * (1 + 1e-6)
    # Laplace heights: CDF 1/2 e^{z/z0} below 0, 1 - 1/2 e^{-z/z0} above
    z = d[:, 2]
    assert ks(z, lambda x: np.where(x < 0, 0.5 * np.exp(x / z0), 1 - 0.5 * np.exp(-x / z0))) > P_MIN
    # uniform azimuth
    phi = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi)
    assert ks(phi / (2 * np.pi), stats.uniform.cdf) > P_MIN
    # circular speed of the exponential disk's enclosed mass, tangential
    vv = v[:nd].astype(np.float64)
    vc = np.sqrt(M * (1 - (1 + R / h) * np.exp(-R / h)) / R)
    assert np.allclose(np.linalg.norm(vv, axis=1), vc, rtol=1e-5)
    assert np.abs((vv[:, 0] * d[:, 0] + vv[:, 1] * d[:, 1]) / (R * vc)).max() < 1e-5
    assert (d[:, 0] * vv[:, 1] - d[:, 1] * vv[:, 0] > 0).all()   # counter-clockwise


def test_reference_scenario_halo_distributions():
    n, h = 200_000, 4.0
    p, _ = ics.reference_scenario(n, seed=3, centre=(1.0, -2.0, 0.5))
    halo = p[n - int(0.2 * n):, :3].astype(np.float64) - [1.0, -2.0, 0.5]
    r = np.linalg.norm(halo, axis=1)
    assert ks(r / (5 * h), stats.uniform.cdf) > P_MIN          # radius uniform in [0, 5h]
    cth = halo[:, 2] / r
    assert ks((cth + 1) / 2, stats.uniform.cdf) > P_MIN       # isotropic: cos theta uniform
    ph = np.mod(np.arctan2(halo[:, 1], halo[:, 0]), 2 * np.pi)
    assert ks(ph / (2 * np.pi), stats.uniform.cdf) > P_MIN


def test_reference_scenario_rotation_and_offsets():
    angles = (0.4, -0.9, 1.3)
    centre, vel0 = (3.0, 1.0, -2.0), (0.1, -0.2, 0.05)
    n = 50_000
    p, v = ics.reference_scenario(n, seed=4, angles=angles, centre=centre, velocity=vel0)
    nd = n - int(0.2 * n)
    rot = ics.rotation_matrix(*angles)
    local = (p[:nd, :3].astype(np.float64) - centre) @ rot      # back into the disk frame
    lv = (v[:nd].astype(np.float64) - vel0) @ rot
    assert np.abs(lv[:, 2]).max() < 1e-5                        # no vertical motion in the frame
    assert np.std(local[:, 2]) < 0.6 < np.std(local[:, 0])      # thin in z, wide in x
    # the disk's angular momentum points along the rotated z axis
    L = np.cross(local, lv).sum(0)
    assert L[2] > 0 and np.hypot(L[0], L[1]) < 1e-3 * L[2]
    assert np.allclose(v[nd:], np.float32(vel0))                 # halo moves with the system


def test_config_disk_distributions():
    n, h, z0 = 100_001, 1.0, 0.1
    p, v = ics.exponential_disk(n, seed=5)
    assert p[0, 3] == np.float32(0.1) and np.allclose(p[0, :3], 0.0)   # the central mass
    assert np.allclose(p[1:, 3].astype(np.float64).sum(), 1.0, rtol=1e-4)
    d = p[1:, :3].astype(np.float64)
    R = np.hypot(d[:, 0], d[:, 1])
    norm = 1.0 - 6.0 * np.exp(-5.0)
    # R = -h ln(u1 u2) (Gamma(2, h)) truncated at 5h
    assert ks(R, lambda r: np.clip((1 - (1 + r / h) * np.exp(-r / h)) / norm, 0, 1)) > P_MIN
    # sech^2(z / z0): CDF (1 + tanh(z / z0)) / 2
    assert ks(d[:, 2], lambda x: 0.5 * (1 + np.tanh(x / z0))) > P_MIN
    # mean tangential speed close to the circular speed of the enclosed mass
    vv = v[1:].astype(np.float64)
    vt = (d[:, 0] * vv[:, 1] - d[:, 1] * vv[:, 0]) / R
    assert 0.9 < np.median(vt / np.linalg.norm(vv, axis=1)) <= 1.0


def test_config_plummer_distributions():
    n = 200_000
    p, v = ics.plummer(n, seed=6)
    x = p[:, :3].astype(np.float64)
    u = v.astype(np.float64)
    assert np.abs(x.mean(0)).max() < 1e-6 and np.abs(u.mean(0)).max() < 1e-6   # COM frame
    r = np.linalg.norm(x, axis=1)
    # M(<r) = r^3 / (1 + r^2)^{3/2}, truncated at 20 a
    m20 = 20.0 ** 3 / (1 + 400.0) ** 1.5
    assert ks(r, lambda s: np.clip(s ** 3 / (1 + s * s) ** 1.5 / m20, 0, 1)) > P_MIN
    # speeds: q = v / v_esc with density q^2 (1 - q^2)^{7/2}
    q = np.linalg.norm(u, axis=1) / (np.sqrt(2.0) * (1 + r * r) ** -0.25)
    grid = np.linspace(0, 1, 2001)
    dens = grid ** 2 * (1 - grid ** 2) ** 3.5
    cdf = integrate.cumulative_trapezoid(dens, grid, initial=0.0)
    cdf /= cdf[-1]
    assert ks(q, lambda s: np.interp(s, grid, cdf)) > P_MIN
    # virial equilibrium: 2K / |W| = 1 with W = -3 pi / 32 (G = M = a = 1)
    K = 0.5 * (p[:, 3] * (u ** 2).sum(1)).sum()
    assert abs(2 * K / (3 * np.pi / 32) - 1) < 0.02


def test_config_merger_orbit():
    k = 200_000
    p, v = ics.disk_merger(k, seed=7)
    m = p[:, 3].astype(np.float64)
    a, b = slice(0, k), slice(k, 2 * k)
    com = lambda s, arr: (m[s, None] * arr[s]).sum(0) / m[s].sum()
    x = p[:, :3].astype(np.float64)
    u = v.astype(np.float64)
    d = com(b, x) - com(a, x)
    w = com(b, u) - com(a, u)
    mt = m.sum()
    assert abs(np.linalg.norm(d) - 20.0) < 0.2
    # parabolic: relative orbital energy 0 up to the sampling noise of each
    # disk's own mean velocity (~ v_c / sqrt(k)); pericentre ~2 from L
    e = 0.5 * (w @ w) - mt / np.linalg.norm(d)
    assert abs(e) < 0.03 * mt / np.linalg.norm(d)
    L = np.linalg.norm(np.cross(d, w))
    assert abs(L ** 2 / (2 * mt) - 2.0) < 0.1
    # the whole system at rest at the origin
    assert np.abs((m[:, None] * x).sum(0) / mt).max() < 0.05
    assert np.abs((m[:, None] * u).sum(0) / mt).max() < 0.01
