"""The seeded initial-condition generators (host code, no GPU) against the distributions they
promise: two exponential disks (scale length 1, height 0.1, log-normal masses sigma 0.5,
parabolic orbit), the Plummer sphere with a uniform background cube and power-law masses
(slope -2.35 on [0.1, 10]), and the log-uniform batch sizes.  These replace the reference's
unseeded boxes (Renderer.swift:441-493); each check is a fit or a goodness-of-fit test that a
wrong parameter fails."""
import numpy as np
import pytest
from scipy import stats


@pytest.fixture(scope="module")
def disks(astro):
    pos, vel, _ = astro.ic_two_disks(100000, 2024)
    x = pos["position"][:, :3].astype(np.float64)
    v = vel["velocity"][:, :3].astype(np.float64)
    m = pos["mass"].astype(np.float64)
    return x, v, m


def halves(a, n=100000):
    return a[:n], a[n:]


def com(x, m):
    return (x * m[:, None]).sum(0) / m.sum()


def disk_frame(x, m, inclined):
    """Positions relative to the disk's centre of mass, rotated back into its own plane."""
    d = x - com(x, m)
    if inclined:  # disk 2 is inclined by 45 degrees about the x axis
        c = s = np.sqrt(0.5)
        y, z = d[:, 1].copy(), d[:, 2].copy()
        d[:, 1], d[:, 2] = c * y + s * z, -s * y + c * z
    return d


def test_two_disks_profiles(disks):
    x, v, m = disks
    for k, (xd, md) in enumerate(zip(halves(x), halves(m))):
        d = disk_frame(xd, md, inclined=k == 1)
        R = np.hypot(d[:, 0], d[:, 1])
        # surface density R e^(-R/Rd): Gamma(2, Rd), so the MLE of Rd is mean(R) / 2
        assert abs(R.mean() / 2 - 1.0) < 0.01, R.mean() / 2
        assert stats.kstest(R, stats.gamma(2, scale=1.0).cdf).pvalue > 1e-3
        # vertical profile exponential in |z| with height 0.1
        assert abs(np.abs(d[:, 2]).mean() - 0.1) < 0.002
        assert stats.kstest(np.abs(d[:, 2]), stats.expon(scale=0.1).cdf).pvalue > 1e-3


def test_two_disks_masses(disks):
    _, _, m = disks
    for md in halves(m):
        assert abs(md.sum() - 0.5) < 1e-5  # normalised to one half per disk
        lm = np.log(md)
        assert abs(lm.std() - 0.5) < 0.005  # log-normal sigma 0.5
        assert stats.kstest((lm - lm.mean()) / 0.5, "norm").pvalue > 1e-3


def test_two_disks_parabolic_orbit(disks):
    # the two centres of mass: separation 10, zero orbital energy (parabolic), pericentre 1
    x, v, m = disks
    (x1, x2), (v1, v2), (m1, m2) = halves(x), halves(v), halves(m)
    dx = com(x1, m1) - com(x2, m2)
    dv = com(v1, m1) - com(v2, m2)
    M = m1.sum() + m2.sum()
    r = np.linalg.norm(dx)
    assert abs(r - 10.0) < 1e-3
    energy = 0.5 * dv @ dv - M / r  # per unit reduced mass, G = 1
    assert abs(energy) < 1e-4 * (M / r)
    L = np.linalg.norm(np.cross(dx, dv))
    assert abs(L * L / (2 * M) - 1.0) < 1e-3  # pericentre of a parabola: L^2 / (2 G M)
    assert dx @ dv < 0  # approaching


def test_two_disks_rotation(disks):
    # circular rotation in each disk's plane: velocities relative to the disk's centre of mass
    # are tangential, counter-clockwise, with the speed of the softened enclosed mass
    x, v, m = disks
    for k, (xd, vd, md) in enumerate(zip(halves(x), halves(v), halves(m))):
        d = disk_frame(xd, md, inclined=k == 1)
        u = disk_frame(vd, md, inclined=k == 1)
        R = np.hypot(d[:, 0], d[:, 1])
        rad = (u[:, 0] * d[:, 0] + u[:, 1] * d[:, 1]) / R
        tan = (d[:, 0] * u[:, 1] - d[:, 1] * u[:, 0]) / R
        o = np.argsort(R)
        menc = np.cumsum(md[o])
        Rs = R[o]
        vc = np.sqrt(menc * Rs * Rs / (Rs * Rs + 0.01) ** 1.5)
        sel = Rs > 0.5  # recentring moved the centre by ~1e-3: small radii see it most
        assert np.abs(u[:, 2]).max() < 1e-6
        assert np.median(np.abs(rad[o][sel]) / vc[sel]) < 1e-2
        rel = (tan[o] - vc) / vc
        assert np.abs(rel[sel]).max() < 0.02 and abs(np.median(rel[sel])) < 2e-3


@pytest.fixture(scope="module")
def plummer_bg(astro):
    n = 200000
    pos, vel, _ = astro.ic_plummer_background(n, 77)
    return n, pos["position"][:, :3].astype(np.float64), vel["velocity"][:, :3].astype(
        np.float64), pos["mass"].astype(np.float64)


def test_plummer_background_cube(plummer_bg):
    n, x, v, _ = plummer_bg
    nbg = int(np.floor(n * 0.1 + 0.5))
    bg = x[n - nbg:]
    assert np.abs(bg).max() <= 10.0 and not v[n - nbg:].any()  # at rest, inside the cube
    for k in range(3):  # uniform on [-10, 10] per axis
        assert stats.kstest(bg[:, k], stats.uniform(loc=-10, scale=20).cdf).pvalue > 1e-3
    # the Plummer part: mass within r is r^3 / (1 + r^2)^(3/2), truncated at r = 10
    r = np.linalg.norm(x[:n - nbg], axis=1)
    assert r.max() < 10.5
    cdf = lambda q: np.minimum(q ** 3 / (1 + q * q) ** 1.5 / (1000 / 101 ** 1.5), 1.0)  # noqa
    assert stats.kstest(r, cdf).pvalue > 1e-3


def test_plummer_background_power_law_masses(plummer_bg):
    # masses m^-2.35 on [0.1, 10], normalised to total 1: the bounds' ratio is 100, and the
    # rescaled masses follow the truncated power law (a wrong slope fails the KS test)
    _, _, _, m = plummer_bg
    assert abs(m.sum() - 1.0) < 1e-5
    scale = 0.1 / m.min()  # the lower bound is reached to within a fraction of a percent
    mm = m * scale
    assert mm.max() <= 10.0 * 1.01 and mm.max() / mm.min() <= 100.0 * (1 + 1e-6)
    k = 1.0 - 2.35
    a, b = 0.1 ** k, 10.0 ** k
    cdf = lambda q: np.clip((np.clip(q, 0.1, 10.0) ** k - a) / (b - a), 0, 1)  # noqa: E731
    assert stats.kstest(mm, cdf).pvalue > 1e-3
    for wrong in (2.0, 2.7):  # a slope off by 0.35 is rejected
        kw = 1.0 - wrong
        aw, bw = 0.1 ** kw, 10.0 ** kw
        assert stats.kstest(mm, lambda q: np.clip((np.clip(q, 0.1, 10.0) ** kw - aw) / (bw - aw),
                                                  0, 1)).pvalue < 1e-6


def test_batch_sizes_log_uniform(astro):
    s = astro.ic_batch_sizes(4096, 1024, 65536, 7).astype(np.float64)
    assert s.min() >= 1024 and s.max() <= 65536
    lo, hi = np.log(1024), np.log(65536)
    assert stats.kstest(np.log(s), stats.uniform(loc=lo, scale=hi - lo).cdf).pvalue > 1e-3
