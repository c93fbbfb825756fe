"""Summarize and sort with no host wait inside a step.

Each level launch reads its cell count on the device; the host only sizes the
grids from the level sizes of an earlier tree, and one single-block launch
takes the deepest levels, including every level deeper than the host knows.
These tests drive a tree that gets ten levels deeper in one step and then
shallower again:
  * each step against the oracle, on the state handed over before it
    (sorted order identical, cell multiset bit-exact, state within 1e-6);
  * many steps in one call against the same steps one call at a time
    (the level sizes the host uses differ, the results are bit-identical),
    in plain mode and with the second moments of the coarse traversal.
"""
import numpy as np
import pytest

from helpers import cell_multiset, make_oracle, make_sim

pytestmark = pytest.mark.gpu

PHYS = dict(theta=0.5, softening=0.05, dt=1e-3)
STEPS = 12


def _closing_pair(n=16384, seed=1):
    """A uniform ball plus two light bodies near its centre on a collision
    course: they share no cell until step 8, where they are 2.5e-6 apart and
    the tree is 20 levels deep instead of 10; afterwards it is 9-10 again."""
    from nbody_amd import universe as U

    st = [np.array(a, dtype=np.float32, copy=True) for a in U.uniform_ball(n, seed=seed)]
    x, y, z, vx, vy, vz, m = st
    x[0], y[0], z[0] = -0.02, 0.001, 0.0
    x[1], y[1], z[1] = 0.02, 0.001001, 0.0
    v = 0.04 / (8.0005 * PHYS["dt"]) / 2
    vx[0], vy[0], vz[0] = v, 0.0, 0.0
    vx[1], vy[1], vz[1] = -v, 0.0, 0.0
    m[0] = m[1] = m[2] * 1e-3
    return st


def test_deepening_tree_tracks_oracle(nb, oracle):
    st = _closing_pair()
    sim = make_sim(st, **PHYS)
    orc = make_oracle(st, 0.5, 0.05, 1e-3, 16)
    depths = []
    for k in range(STEPS):
        x, y, z, vx, vy, vz = sim.state()
        acc = sim.accel()
        orc.set_state(x, y, z, vx, vy, vz, st[6])
        orc.set_accel(acc[:, 0], acc[:, 1], acc[:, 2], k - 1)
        sim.steps(1)
        assert orc.step(1) == 0
        _, od = orc.paths()
        depths.append(int(od.max()))
        assert np.array_equal(sim.sorted(), orc.sorted()), f"step {k}: sorted order differs"
        gc, oc = sim.cells(), orc.cells()
        g = cell_multiset(gc["count"], gc["x"], gc["y"], gc["z"], gc["m"])
        o = cell_multiset(oc["count"], oc["x"], oc["y"], oc["z"], oc["m"])
        assert len(g) == len(o) and np.array_equal(g, o), f"step {k}: cells differ"
        gs = np.stack(sim.state(), axis=1)
        os_ = np.stack(orc.state(), axis=1)
        d = np.abs(gs.astype(np.float64) - os_) / np.maximum(np.abs(os_), 1e-3)
        assert d.max() < 1e-6, f"step {k}: max rel state diff {d.max():.3e}"
    # the workload does what it is for: ten levels beyond the previous tree, then fewer
    assert max(depths) >= depths[0] + 8, depths
    assert depths[-1] < max(depths), depths
    sim.close()


@pytest.mark.parametrize("mode", ["plain", "coarse"])
def test_queued_steps_match_single_steps(nb, mode):
    from nbody_amd import _lib

    kw = dict(PHYS, flags=_lib.NB_FP64_POSITIONS)
    if mode == "coarse":
        kw.update(method=_lib.NB_MODE_COARSE, coarse_theta=0.9, coarse_alpha=0.02, vote_width=64)
    st = _closing_pair()
    a = make_sim(st, **kw)
    b = make_sim(st, **kw)
    a.steps(STEPS)
    for _ in range(STEPS):
        b.steps(1)
    for ga, gb in zip(a.state(), b.state()):
        assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
    assert np.array_equal(a.positions64(), b.positions64())
    assert np.array_equal(a.sorted(), b.sorted())
    assert a.counters().max_depth == b.counters().max_depth >= 18
    a.close()
    b.close()
