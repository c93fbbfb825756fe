"""Initial conditions: determinism, shapes and the physical properties each
generator promises (CPU only)."""
import numpy as np
import pytest


@pytest.mark.parametrize("name", ["uniform_ball", "plummer", "two_plummer", "gaussian_clusters", "halo_disk"])
def test_seeded_and_float32(nb, name):
    fn = nb.universe.CONFIGS[name]
    a = fn(5000)
    b = fn(5000)
    c = fn(5000, seed=99)
    assert len(a) == 7 and all(x.dtype == np.float32 and x.shape == (5000,) for x in a)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    assert np.isclose(a[6].astype(np.float64).sum(), 1.0, rtol=1e-5)
    assert all(np.all(np.isfinite(x)) for x in a)


def test_uniform_ball(nb):
    x, y, z, vx, vy, vz, m = nb.universe.uniform_ball(200_000, seed=1)
    r = np.sqrt(x.astype(float) ** 2 + y ** 2 + z ** 2)
    assert r.max() <= 1.0 + 1e-6
    assert abs(np.median(r) - 0.5 ** (1 / 3)) < 0.01  # uniform in volume
    # solid-body rotation about z with omega = 0.5 plus N(0, 0.01^2) noise
    res = np.stack([vx + 0.5 * y, vy - 0.5 * x, vz], 1)
    assert abs(res.std() - 0.01) < 5e-4


def test_plummer_profile(nb):
    """Plummer (a = 1, M = 1): half-mass radius 1.3048 a, kinetic energy
    3 pi / 64 (virial equilibrium), truncated at X = 0.999 of the mass."""
    x, y, z, vx, vy, vz, m = nb.universe.plummer(400_000, seed=2)
    r = np.sqrt(x.astype(float) ** 2 + y ** 2 + z ** 2)
    assert abs(np.median(r) - 1.3048) < 0.02
    ke = 0.5 * np.sum(m * (vx.astype(float) ** 2 + vy ** 2 + vz ** 2))
    assert abs(ke - 3 * np.pi / 64) / (3 * np.pi / 64) < 0.03


def test_two_plummer(nb):
    x, y, z, vx, vy, vz, m = nb.universe.two_plummer(100_000, seed=3)
    left = x < 0
    assert abs(left.mean() - 0.5) < 0.01
    assert abs(np.median(x[left]) + 5) < 0.1 and abs(np.median(x[~left]) - 5) < 0.1
    assert abs(vx[left].mean() - 0.3) < 0.02 and abs(vx[~left].mean() + 0.3) < 0.02


def test_gaussian_clusters(nb):
    x, y, z, vx, vy, vz, m = nb.universe.gaussian_clusters(200_000, seed=4)
    assert np.all(np.abs(np.stack([x, y, z])) < 12)
    assert np.allclose(m, 1 / 200_000)


def test_halo_disk(nb):
    x, y, z, vx, vy, vz, m = nb.universe.halo_disk(80_000, seed=5)
    disk = slice(60_000, None)
    R = np.sqrt(x[disk].astype(float) ** 2 + y[disk] ** 2)
    assert abs(np.mean(R) - 6.0) < 0.2          # gamma(2, 3) has mean 6
    assert abs(np.mean(np.abs(z[disk])) - 0.3) < 0.02
    # circular: velocity perpendicular to the radius
    dot = x[disk] * vx[disk] + y[disk] * vy[disk]
    assert np.max(np.abs(dot)) < 1e-4


def test_reference_ports(nb):
    U = nb.universe
    x, y, z, vx, vy, vz, m = U.random_cubic(10_000, 6.0, seed=1)
    assert np.all(np.abs(np.stack([x, y, z])) <= 3.0) and np.all(vx == 0)
    assert np.allclose(m, 1 / 10_000)
    x, y, z, *_ = U.spherical(10_000)
    assert np.allclose(x.astype(float) ** 2 + y ** 2 + z ** 2, 1.0, atol=1e-6)
    x, y, z, *_ = U.monte_carlo_spherical(10_000)
    assert np.allclose(x.astype(float) ** 2 + y ** 2 + z ** 2, 1.0, atol=1e-6)
    x, y, z, *_ = U.lon_lat_spherical(1000)
    assert np.allclose(np.sqrt(x.astype(float) ** 2 + y ** 2 + z ** 2), 5.0, atol=1e-5)
    st = U.rotating_disk(2048)
    assert st[6][0] == 1.0 and st[0][0] == 0.0
    r = np.sqrt(st[0][1:].astype(float) ** 2 + st[1][1:] ** 2)
    assert r.min() >= 0.05 and r.max() <= 3.55
    st = U.plummer_henon(10_000)
    assert all(np.all(np.isfinite(a)) for a in st)
    assert U.eight_body()[0][0] == np.float32(-1.1)
    assert U.big_tree()[6][3] == 0.0


def test_generator_protocol(nb):
    U = nb.universe
    st = U.uniform_ball(100)
    g = U.ArrayUniverse(st)
    arrs = [np.zeros(120, np.float32) for _ in range(7)]
    g.generate(10, 100, *arrs)
    assert np.array_equal(arrs[0][10:110], st[0]) and arrs[0][0] == 0
    with pytest.raises(ValueError):
        g.generate(0, 50, *arrs)
    f = U.FunctionUniverse(U.uniform_ball, seed=1)
    f.generate(0, 100, *arrs)
    assert np.array_equal(arrs[0][:100], st[0])
