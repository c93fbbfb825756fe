"""GPU parity: each stage of libnbody.so against the CPU oracle on the same
seeded inputs (sizes the oracle finishes in seconds).

Bars (bit-exact where the arithmetic is fixed, stated tolerance otherwise):
  bbox        root centre and radius bit-exact; the reference test's own
              |root - 0.5 (min + max)| < 1e-5 (TestBoundingBox.java:236-251)
  build       every body's octant path and depth bit-exact
  summarize   the multiset of (count, centre of mass, mass) of the cells
              bit-exact (FP64 accumulation in child-slot order on both sides)
  sort        sorted[] identical
  force       per-body relative acceleration error <= 1e-5 at the same theta
              and vote width (FP32 Kahan-compensated traversal against the
              oracle's FP64 accumulation)
  step        positions and velocities after integrate within 1e-6 relative
"""
import numpy as np
import pytest

from helpers import cell_multiset, make_oracle, make_sim, rel_err, tight_clump

pytestmark = pytest.mark.gpu

ACC_TOL = 1e-5


def _universes():
    from nbody_amd import universe as U

    return {
        "uniform_ball_16k": lambda: U.uniform_ball(16384, seed=1),
        "plummer_32k": lambda: U.plummer(32768, seed=2),
        "clusters_32k": lambda: U.gaussian_clusters(32768, seed=4),
        "ragged_1000": lambda: U.uniform_ball(1000, seed=7),
        "ragged_12345": lambda: U.plummer(12345, seed=8),
        # 27 bodies sharing their first 21 octants (helpers.tight_clump)
        "tight_run_8219": lambda: tight_clump(),
    }


CASES = ["uniform_ball_16k", "plummer_32k", "clusters_32k", "ragged_1000", "ragged_12345", "tight_run_8219"]
PHYS = dict(theta=0.5, softening=0.05, dt=1e-3)


def _run_stages(sim, orc, last):
    for st in range(1, last + 1):
        sim.run_stage(st)
    assert orc.run_stages(1, last) == 0


@pytest.mark.parametrize("case", CASES)
def test_bbox_bitexact(nb, oracle, case):
    st = _universes()[case]()
    sim = make_sim(st, **PHYS)
    orc = make_oracle(st, 0.5, 0.05, 1e-3, 16)
    _run_stages(sim, orc, 1)
    g, o = sim.root(), orc.root()
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), (g, o)
    # the reference's own assertion (TestBoundingBox.java:236-251)
    x, y, z = st[0], st[1], st[2]
    for k, a in enumerate((x, y, z)):
        assert abs(np.float32(0.5) * (a.min() + a.max()) - g[k]) < 1e-5
    sim.close()


def test_bbox_reference_test_universe(nb, oracle):
    """TestBoundingBox.init: 32768 bodies from RandomCubicUniverseGenerator(32768)
    (TestBoundingBox.java:93-94), only the bounding-box kernel runs (:207)."""
    from nbody_amd import universe as U

    st = U.random_cubic(32768, 32768.0, seed=11)
    sim = make_sim(st, **PHYS)
    sim.run_stage(1)
    g = sim.root()
    for k in range(3):
        a = st[k]
        assert abs(np.float32(0.5) * (a.min() + a.max()) - g[k]) < 1e-5
    sim.close()


@pytest.mark.parametrize("case", CASES)
def test_membership_bitexact(nb, oracle, case):
    st = _universes()[case]()
    sim = make_sim(st, **PHYS)
    orc = make_oracle(st, 0.5, 0.05, 1e-3, 16)
    _run_stages(sim, orc, 2)
    gp, gd = sim.paths()
    op, od = orc.paths()
    assert np.all(gd > 0), "a body is not reachable by descent"
    assert np.array_equal(gd, od)
    assert np.array_equal(gp, op)
    assert sim.counters().max_depth == orc.scalars()["max_depth"]
    sim.close()


@pytest.mark.parametrize("case", CASES)
def test_summarize_and_sort_bitexact(nb, oracle, case):
    st = _universes()[case]()
    sim = make_sim(st, **PHYS)
    orc = make_oracle(st, 0.5, 0.05, 1e-3, 16)
    _run_stages(sim, orc, 4)
    gc = sim.cells()
    oc = orc.cells()
    g = cell_multiset(gc["count"], gc["x"], gc["y"], gc["z"], gc["m"])
    o = cell_multiset(oc["count"], oc["x"], oc["y"], oc["z"], oc["m"])
    assert len(g) == len(o)
    assert np.array_equal(g, o)
    assert gc["count"][0] == len(st[0])  # root holds every body
    gs, os_ = sim.sorted(), orc.sorted()
    assert np.array_equal(np.sort(gs), np.arange(len(st[0])))
    assert np.array_equal(gs, os_)
    sim.close()


def _collapse(n=32768, seed=12):
    """n - 8 bodies uniform in [-1, 1]^3 heading for a ball of radius 0.01
    around (0.3, 0.3, 0.3), which they reach after a step of dt 1 (masses 1e-9:
    gravity moves them by less than 1e-9), and 8 bodies at rest on the corners,
    which keep the bounding box."""
    g = np.random.default_rng(seed)
    p = g.uniform(-1.0, 1.0, (n, 3))
    p[:8] = [[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
    target = 0.3 + g.uniform(-0.01, 0.01, (n, 3))
    v = target - p
    v[:8] = 0.0
    m = np.full(n, 1e-9)
    return tuple(np.ascontiguousarray(a, np.float32) for a in (p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], m))


@pytest.mark.parametrize("case,dt,large,fused", [("plummer_32k", 0.0, False, False),
                                                 ("plummer_32k", 1e-3, False, False),
                                                 ("plummer_32k", 1e-3, False, True),
                                                 ("collapse_32k", 1.0, True, False),
                                                 ("collapse_32k", 1.0, True, True)])
def test_warm_build_tracks_oracle(nb, oracle, case, dt, large, fused):
    """After a step the build starts from that tree's body order (sorted[]):
    bucketed by splitters taken from that tree's sorted keys, each bucket
    sorted on chip (or through device memory when it is large). The second
    tree against the oracle's tree of the same positions: membership, cells
    and sorted[] bit for bit. dt 0 leaves the positions (and the order)
    unchanged; in the collapse case nearly every body moves into a ball that
    one bucket of the last tree covers, which is then far too large for LDS.
    With the fused update the build reads the positions from the records that
    update wrote in the last order instead of gathering them."""
    from nbody_amd import _lib

    st = _collapse() if case == "collapse_32k" else _universes()[case]()
    sim = make_sim(st, theta=0.5, softening=0.05, dt=dt, flags=_lib.NB_FUSED_UPDATE if fused else 0)
    sim.steps(1)
    x, y, z, vx, vy, vz = sim.state()
    orc = make_oracle((x, y, z, vx, vy, vz, st[6]), 0.5, 0.05, dt, 16)
    before = sim.counters().large_buckets
    for stage in (1, 2):
        sim.run_stage(stage)
    assert orc.run_stages(1, 2) == 0
    gp, gd = sim.paths()
    op, od = orc.paths()
    assert np.all(gd > 0)
    assert np.array_equal(gd, od)
    assert np.array_equal(gp, op)
    for stage in (3, 4):
        sim.run_stage(stage)
    assert orc.run_stages(3, 4) == 0
    gc, oc = sim.cells(), orc.cells()
    assert np.array_equal(cell_multiset(gc["count"], gc["x"], gc["y"], gc["z"], gc["m"]),
                          cell_multiset(oc["count"], oc["x"], oc["y"], oc["z"], oc["m"]))
    assert np.array_equal(sim.sorted(), orc.sorted())
    grew = sim.counters().large_buckets - before
    assert (grew > 0) == large, grew
    sim.close()


@pytest.mark.parametrize("vote", [16, 32, 64])
@pytest.mark.parametrize("case", CASES)
def test_force_parity(nb, oracle, case, vote):
    st = _universes()[case]()
    sim = make_sim(st, vote_width=vote, **PHYS)
    orc = make_oracle(st, 0.5, 0.05, 1e-3, vote)
    _run_stages(sim, orc, 5)
    ag, ao = sim.accel(), orc.accel()
    err = rel_err(ag, ao)
    assert np.all(np.isfinite(ag))
    assert err.max() <= ACC_TOL, f"max rel err {err.max():.3e} at body {err.argmax()}"
    assert sim.counters().interactions == orc.scalars()["interactions"]
    sim.close()


@pytest.mark.parametrize("fused", [False, True])
def test_full_steps_track_oracle(nb, oracle, fused):
    """Three steps; before each step the oracle is handed the engine's state,
    so every step is compared on identical inputs (first step: no velocity
    correction, calculateforce.cl:174)."""
    from nbody_amd import _lib

    st = _universes()["uniform_ball_16k"]()
    flags = _lib.NB_FUSED_UPDATE if fused else 0
    sim = make_sim(st, flags=flags, **PHYS)
    orc = make_oracle(st, 0.5, 0.05, 1e-3, 16)
    for k in range(3):
        x, y, z, vx, vy, vz = sim.state()
        acc = sim.accel()
        orc.set_state(x, y, z, vx, vy, vz, st[6])
        orc.set_accel(acc[:, 0], acc[:, 1], acc[:, 2], k - 1)
        sim.steps(1)
        assert orc.step(1) == 0
        g = np.stack(sim.state(), axis=1)
        o = np.stack(orc.state(), axis=1)
        assert sim.counters().step == k
        d = np.abs(g.astype(np.float64) - o) / np.maximum(np.abs(o), 1e-3)
        assert d.max() < 1e-6, f"step {k}: max rel state diff {d.max():.3e}"
    sim.close()


def test_fp64_positions_track_oracle(nb, oracle):
    from nbody_amd import _lib
    import pyoracle

    st = _universes()["plummer_32k"]()
    sim = make_sim(st, flags=_lib.NB_FP64_POSITIONS, **PHYS)
    orc = make_oracle(st, 0.5, 0.05, 1e-3, 16, flags=pyoracle.FP64_POS)
    sim.steps(1)
    assert orc.step(1) == 0
    g64, o64 = sim.positions64(), orc.positions64()
    assert np.max(np.abs(g64 - o64)) < 1e-9
    sim.close()


def test_known_trees(nb):
    """Hand-derived trees of the reference's test universes."""
    from nbody_amd import universe as U

    # EightBodyUniverse: root (-0.05, 0.05, 0), radius 1.1, one level
    sim = make_sim(U.eight_body(), **PHYS)
    sim.run_stage(1)
    sim.run_stage(2)
    root = sim.root()
    assert np.allclose(root, [-0.05, 0.05, 0.0, 1.1], atol=1e-6)
    p, d = sim.paths()
    assert list(d) == [1] * 8
    assert list(p[:, 0]) == [0, 1, 2, 4, 3, 6, 5, 7]
    sim.close()
    # BigTreeUniverse: bodies 0 and 3 share octant 0 and split below it
    sim = make_sim(U.big_tree(), **PHYS)
    sim.run_stage(1)
    sim.run_stage(2)
    p, d = sim.paths()
    assert d[1] == 1 and d[2] == 1 and p[1, 0] == 1 and p[2, 0] == 7
    assert d[0] >= 2 and d[3] >= 2 and (p[0, 0] & 7) == 0 and (p[3, 0] & 7) == 0
    sim.close()


@pytest.mark.parametrize("fused", [False, True])
def test_single_body(nb, oracle, fused):
    """N = 1: the root holds the body, which meets only itself (d = 0, so no
    force: calculateforce.cl:138-151); it drifts at constant velocity.
    Three steps against the oracle bit for bit and against x0 + 3 v dt."""
    from nbody_amd import _lib

    st = tuple(np.array([v], np.float32) for v in (0.25, -0.5, 1.0, 0.3, -0.2, 0.1, 2.0))
    sim = make_sim(st, flags=_lib.NB_FUSED_UPDATE if fused else 0, **PHYS)
    orc = make_oracle(st, 0.5, 0.05, 1e-3, 16)
    sim.steps(3)
    assert orc.step(3) == 0
    assert np.array_equal(sim.accel(), np.zeros((1, 3), np.float32))
    g = np.array(sim.state(), np.float32)
    o = np.array(orc.state(), np.float32)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), (g, o)
    x0 = np.array(st[:3], np.float64).ravel()
    v0 = np.array(st[3:6], np.float64).ravel()
    assert np.allclose(g[:3].ravel(), x0 + 3 * 1e-3 * v0, rtol=0, atol=1e-6)
    assert np.array_equal(g[3:].ravel(), np.array(st[3:6], np.float32).ravel())
    assert sim.counters().step == 2
    sim.close()


def test_coincident_bodies_report_error(nb):
    """Two bodies at identical FP32 coordinates never separate (the reference
    exhausts its cell pool, buildtree.cl:112-120): the step must fail with a
    status code, not hang."""
    from nbody_amd import NbError, universe as U

    st = list(U.uniform_ball(4096, seed=3))
    for a in st[:3]:
        a[17] = a[5]
    sim = make_sim(tuple(st), **PHYS)
    with pytest.raises(NbError) as e:
        sim.steps(1)
    assert e.value.code in (-2, -3)
    sim.close()


@pytest.mark.parametrize("case", ["clusters_32k", "ragged_12345"])
def test_direct_sum_matches_oracle(nb, oracle, case):
    """The device FP64 direct sum (the error yardstick) against the oracle's.
    12,345 bodies split into 64 source chunks of 193 (the last one 186):
    every tile is partial, so the source loop's rounds of four and its
    remainder both run."""
    st = _universes()[case]()
    n = len(st[0])
    sim = make_sim(st, **PHYS)
    idx = np.arange(0, n, 97, dtype=np.int32)
    g = sim.direct_accel(idx)
    o = oracle.direct_fp64(st[0], st[1], st[2], st[6], 0.05, idx)
    assert np.max(np.abs(g - o) / np.linalg.norm(o, axis=1)[:, None]) < 1e-12
    sim.close()


@pytest.mark.parametrize("vote", [16, 64])
def test_probe_and_error(nb, oracle, vote):
    """The probe pass gives the accelerations of a force stage on the same
    positions without advancing the bodies; theta -> 0 has only the FP32
    rounding error against the FP64 direct sum."""
    st = _universes()["plummer_32k"]()
    sim = make_sim(st, vote_width=vote, **PHYS)
    probe = sim.probe_accel()
    assert sim.counters().step == -1
    assert np.array_equal(sim.state()[0], st[0])
    for s in range(1, 6):
        sim.run_stage(s)
    assert np.array_equal(probe, sim.accel())
    sim.close()
    sim = make_sim(st, vote_width=vote, theta=1e-4, softening=0.05, dt=1e-3)
    e = sim.force_error(512, seed=3)
    assert e["max"] < 2e-5 and e["median"] < 2e-6
    sim.close()
    sim = make_sim(st, vote_width=vote, theta=0.7, softening=0.05, dt=1e-3)
    e = sim.force_error(2048, seed=3)
    assert 1e-5 < e["median"] < 1e-2
    sim.close()
