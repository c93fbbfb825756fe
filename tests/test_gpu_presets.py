"""The reference's universe generators, seeded ports in nbody_amd.universe,
run through the engine at the reference's shipped constants (THETA 1.5,
EPSILON 0.0025, TIMESTEP 0.025, work-group vote of 16 lanes:
calculateforce.cl:11-18, GPUBarnesHutNBodySimulation.java:47) against the
oracle, three steps each with the oracle handed the engine's state before
every step.

Bars: every body's octant path and depth bit-exact, the same interaction
count, accelerations within 1e-5 relative, and the step's velocities and
positions within what that acceleration bar allows through one kick-drift-
kick at dt = 0.025 (|dv| <= dt 1e-5 |a|, |dx| <= dt |dv|) plus FP32 rounding. Presets: the visualizer default
RotatingDiskGalaxyGenerator(3.5f, 1, 1) at 2048 * 16 bodies
(NBodyVisualizer.java:212; a central body of mass 1 among bodies of mass
1/N), PlummerUniverseGenerator (Henon units), RandomCubicUniverseGenerator,
SphericalUniverseGenerator, MonteCarloSphericalUniverseGenerator and
LonLatSphericalUniverseGenerator.
"""
import math

import numpy as np
import pytest

from helpers import make_oracle, make_sim, rel_err

pytestmark = pytest.mark.gpu

THETA = math.sqrt(1.5)  # the reference's THETA is theta^2
SOFT = 0.05             # EPSILON = softening^2
DT = 0.025


def _presets():
    from nbody_amd import universe as U

    return {
        "rotating_disk_32768": lambda: U.rotating_disk(2048 * 16, 3.5, 1.0, 1.0, seed=1),
        "plummer_henon_32768": lambda: U.plummer_henon(32768, seed=2),
        "random_cubic_8192": lambda: U.random_cubic(8192, 10.0, seed=3),
        "spherical_8192": lambda: U.spherical(8192, seed=4),
        "monte_carlo_spherical_8192": lambda: U.monte_carlo_spherical(8192, seed=5),
        "lon_lat_spherical_8192": lambda: U.lon_lat_spherical(8192, seed=6),
    }


@pytest.mark.parametrize("preset", list(_presets()))
def test_preset_tracks_oracle(nb, oracle, preset):
    st = _presets()[preset]()
    n = len(st[0])
    sim = make_sim(st, theta=THETA, softening=SOFT, dt=DT, vote_width=16)
    orc = make_oracle(st, THETA, SOFT, DT, 16)
    for k in range(3):
        x, y, z, vx, vy, vz = sim.state()
        acc = sim.accel()
        orc.set_state(x, y, z, vx, vy, vz, st[6])
        orc.set_accel(acc[:, 0], acc[:, 1], acc[:, 2], k - 1)
        # membership right after the build (the oracle's summarize compacts
        # its child slots, as the reference's does, summarizetree.cl:77-81)
        sim.run_stage(1)
        sim.run_stage(2)
        assert orc.run_stages(1, 2) == 0
        gp, gd = sim.paths()
        op, od = orc.paths()
        assert np.all(gd > 0), f"step {k}: a body is not reachable by descent"
        assert np.array_equal(gd, od) and np.array_equal(gp, op), f"step {k}: membership differs"
        for s in range(3, 6):
            sim.run_stage(s)
        assert orc.run_stages(3, 5) == 0
        ag, ao = sim.accel(), orc.accel()
        err = rel_err(ag, ao)
        assert np.all(np.isfinite(ag))
        assert err.max() <= 1e-5, f"step {k}: max rel err {err.max():.3e} at body {err.argmax()}"
        assert sim.counters().interactions == orc.scalars()["interactions"]
        sim.run_stage(6)
        assert orc.run_stages(6, 6) == 0
        # the step carries the acceleration bar through one kick-drift-kick at
        # the reference's dt = 0.025: |dv| <= dt |da| and |dx| <= dt |dv|,
        # with |da| <= 1e-5 |a| (a is up to ~150 next to the disk's central
        # mass), plus FP32 rounding of the state itself
        g = np.stack(sim.state(), axis=1).astype(np.float64)
        o = np.stack(orc.state(), axis=1)
        amag = np.linalg.norm(ao.astype(np.float64), axis=1)[:, None]
        bv = DT * 1e-5 * amag + 1e-6 * np.abs(o[:, 3:])
        bx = DT * bv + 1e-6 * np.abs(o[:, :3])
        assert np.all(np.abs(g[:, 3:] - o[:, 3:]) <= bv + 1e-7), f"step {k}: velocities differ"
        assert np.all(np.abs(g[:, :3] - o[:, :3]) <= bx + 1e-7), f"step {k}: positions differ"
        assert sim.counters().step == k
    assert n == len(sim.state()[0])
    sim.close()
