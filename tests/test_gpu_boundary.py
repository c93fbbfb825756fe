"""GPU tests of the renderer export and of resuming a learned run, through libnbody.so.

Bars:
  copyVertices   nb_copy_vertices writes exactly (x, y, z, 1) and (vx, vy, vz, 1)
                 of the values nb_get_state reads (copyvertices.cl:7-17), bit for
                 bit: with either output pointer NULL, with both NULL (nothing to
                 do), on a ragged body count, after several steps (the staging
                 buffer is reused), and with FP64 positions (the float4 then holds
                 the FP32 rounding of the FP64 position the step kept). The Python
                 mirror's GL_INTEROP step() fills its vertex arrays the same way.
  learned resume a learned run with sampled checks and check_adapt, checkpointed
                 after two failed checks (alpha tightened twice, the check
                 schedule mid-way), resumes bit-identical to the uninterrupted run
                 through its third failed check and the fallback to plain theta;
                 counters and control state included.
  set_state      a new universe on a used handle starts from a new handle's
                 control state (alpha, fallback, check schedule).
"""
import numpy as np
import pytest

from helpers import make_sim, random_model

pytestmark = pytest.mark.gpu


def _expect4(a, b, c):
    return np.stack([a, b, c, np.ones_like(a)], 1).astype(np.float32)


@pytest.mark.parametrize("fp64", [False, True])
def test_copy_vertices_exact(nb, fp64):
    from nbody_amd import universe as U

    n = 5000  # not a multiple of the 256-thread block
    st = U.plummer(n, seed=31)
    flags = (nb._lib.NB_FP64_POSITIONS | nb._lib.NB_FUSED_UPDATE) if fp64 else 0
    sim = make_sim(st, theta=0.5, softening=0.01, dt=1e-3, vote_width=64, flags=flags)
    for k in range(3):
        sim.steps(1)
        x, y, z, vx, vy, vz = sim.state()
        pos = np.full((n, 4), np.nan, np.float32)
        vel = np.full((n, 4), np.nan, np.float32)
        sim.copy_vertices(pos, vel)
        assert np.array_equal(pos.view(np.uint32), _expect4(x, y, z).view(np.uint32)), k
        assert np.array_equal(vel.view(np.uint32), _expect4(vx, vy, vz).view(np.uint32)), k
        if fp64:
            p64 = sim.positions64()
            assert np.array_equal(pos[:, :3], p64.astype(np.float32))
    # one output at a time: the other pointer is NULL and nothing else is written
    pos = np.full((n, 4), np.nan, np.float32)
    sim.copy_vertices(pos, None)
    assert np.array_equal(pos.view(np.uint32), _expect4(x, y, z).view(np.uint32))
    vel = np.full((n, 4), np.nan, np.float32)
    sim.copy_vertices(None, vel)
    assert np.array_equal(vel.view(np.uint32), _expect4(vx, vy, vz).view(np.uint32))
    sim.copy_vertices(None, None)
    sim.close()


def test_gl_interop_step_fills_vertices(nb):
    """The plugin API's GL_INTEROP path (GPUBarnesHutNBodySimulation.java:229-246,
    265-266): after step() the mirror's vertex arrays hold the new state."""
    from nbody_amd import HipBarnesHutNBodySimulation, Mode, universe as U

    n = 4096
    st = U.uniform_ball(n, seed=33)
    sim = HipBarnesHutNBodySimulation(Mode.GL_INTEROP, n, U.ArrayUniverse(st), theta=0.5, softening=0.05,
                                      dt=1e-3, vote_width=16)
    sim.init(None)
    sim.initGLBuffers(None, 1, 2)
    sim.step()
    x, y, z, vx, vy, vz = sim.state()
    assert np.array_equal(sim.positions, _expect4(x, y, z))
    assert np.array_equal(sim.velocities, _expect4(vx, vy, vz))
    sim.close()


def _learned_kw(nb):
    L = nb._lib
    # a tolerance no step meets: every check fails; check_adapt tightens alpha
    # after the first two and the third ends the learned force
    return dict(theta=0.4, softening=0.01, dt=1e-3, vote_width=64, coarse_theta=0.9, coarse_alpha=0.02,
                method=L.NB_MODE_LEARNED, check_tol=1e-12, check_every=2, check_samples=256, check_adapt=1,
                flags=L.NB_FP64_POSITIONS | L.NB_FUSED_UPDATE)


def test_learned_checkpoint_resume_bit_identical(nb, tmp_path):
    from nbody_amd import universe as U

    st = U.plummer(20000, seed=32)
    blob = random_model(6)
    kw = _learned_kw(nb)
    a = make_sim(st, **kw)
    a.set_model(blob)
    a.steps(7)
    ca = a.counters()
    assert ca.checks == 3 and ca.fallbacks == 3 and a.control().fell_back == 1

    b = make_sim(st, **kw)
    b.set_model(blob)
    b.steps(3)
    cb = b.control()
    assert cb.fail_streak == 2 and cb.steps_since_model == 3 and np.isclose(cb.coarse_alpha, 0.02 * 0.25)
    path = str(tmp_path / "learned.npz")
    b.checkpoint(path)

    c = make_sim(st, **kw)  # no model yet: restore installs the checkpoint's
    c.restore(path)
    c.steps(4)
    assert np.array_equal(a.positions64(), c.positions64())
    for u, w in zip(a.state(), c.state()):
        assert np.array_equal(u.view(np.uint32), w.view(np.uint32))
    assert np.array_equal(a.accel().view(np.uint32), c.accel().view(np.uint32))
    cc = c.counters()
    assert (cc.step, cc.checks, cc.fallbacks) == (ca.step, ca.checks, ca.fallbacks)
    assert cc.coarse_alpha == ca.coarse_alpha
    ka, kc = a.control(), c.control()
    for f in ("steps_since_model", "check_seed", "coarse_alpha", "fell_back", "fail_streak"):
        assert getattr(ka, f) == getattr(kc, f), f
    assert np.array_equal(a.get_model(), c.get_model())
    for s in (a, b, c):
        s.close()


def test_set_state_resets_control(nb):
    from nbody_amd import universe as U

    st = U.plummer(20000, seed=34)
    kw = _learned_kw(nb)
    a = make_sim(st, **kw)
    a.set_model(random_model(7))
    fresh = a.control()
    a.steps(5)
    k = a.control()
    assert k.fell_back == 1 and k.coarse_alpha < 0.02
    a.set_state(*st)
    k = a.control()
    for f in ("steps_since_model", "check_seed", "coarse_alpha", "fell_back", "fail_streak"):
        assert getattr(k, f) == getattr(fresh, f), f
    # the counters of the run so far are kept
    assert k.checks == 3 and k.fallbacks == 3
    a.steps(1)
    assert a.control().fell_back == 0 and a.counters().checks == 4
    a.close()
