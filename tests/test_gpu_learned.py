"""GPU tests of the learned coarse-theta correction, through libnbody.so.

Bars:
  coarse traversal  accelerations within 1e-5 relative of the oracle's FP64
                    restatement (same opening decisions, FP32 Kahan sums); the
                    24 features within 1e-3 absolute + 1e-3 relative on all but
                    a handful of bodies (FP32 sums against FP64)
  BF16 inference    the matrix-core network against the oracle's FP64 forward
                    pass on the same normalized inputs: rms difference <= 1% of
                    the rms output, max <= 5% of the max output (BF16 weights and
                    activations, FP32 accumulation)
  training kernels  FP32 gradient within 1e-4 of the largest FP64 component;
                    first Adam step = -lr sign(g) (closed form); four steps
                    against an FP64 Adam restatement driven by oracle gradients
  ablation          after training on the simulation's own pairs, the learned
                    method's median error against the FP64 direct sum is below
                    the coarse traversal's alone on the same state
  fallback          a check tolerance nobody meets makes the step plain theta;
                    with check_adapt alpha tightens instead (exact halvings at
                    such a tolerance) until three failures in a row
"""
import numpy as np
import pytest

from helpers import make_oracle, make_sim, random_model as _random_model, rel_err

pytestmark = pytest.mark.gpu

EPS = 0.01
THETA_C = 0.9
ALPHA = 0.02


def _coarse_sim(st, method=None, **kw):
    import nbody_amd._lib as L

    return make_sim(st, theta=0.4, softening=EPS, dt=1e-3, vote_width=64, coarse_theta=THETA_C,
                    method=L.NB_MODE_COARSE if method is None else method, **kw)


def _hand_over(sim, st, step_offset=0):
    """An oracle holding the engine's current state: positions, velocities,
    the accelerations of its last step (the relative test's a_old) and its
    step counter."""
    x, y, z, vx, vy, vz = sim.state()
    orc = make_oracle((x, y, z, vx, vy, vz, st[6]), 0.4, EPS, 1e-3, 64, flags=1)
    a = sim.accel()
    orc.set_accel(*(np.ascontiguousarray(a[:, j]) for j in range(3)), sim.counters().step + step_offset)
    return orc


def _massless_half(n, seed):
    """A ball whose bodies with x > 0.2 have no mass: the cells holding only
    them have mass 0, their centre at the origin (where the reference's
    summarize gives 0 * inf = NaN) and a relative-test factor of 0."""
    from nbody_amd import universe as U

    st = U.uniform_ball(n, seed=seed)
    m = np.where(st[0] > 0.2, np.float32(0), st[6]).astype(np.float32)
    return tuple(st[:6]) + (m,)


@pytest.mark.parametrize("case", ["plummer_16k", "clusters_32k", "ragged_5000", "clusters_32k_history",
                                  "massless_8k_history"])
def test_coarse_features_match_oracle(nb, oracle, case):
    """The `history` cases run three coarse steps first, so the relative test
    (alpha 0.02, on the previous accelerations) takes part in every decision."""
    from nbody_amd import universe as U

    st = {"plummer_16k": lambda: U.plummer(16384, seed=2),
          "clusters_32k": lambda: U.gaussian_clusters(32768, seed=4),
          "clusters_32k_history": lambda: U.gaussian_clusters(32768, seed=4),
          "massless_8k_history": lambda: _massless_half(8192, 5),
          "ragged_5000": lambda: U.uniform_ball(5000, seed=9)}[case]()
    sim = _coarse_sim(st, coarse_alpha=ALPHA)
    if case.endswith("history"):
        sim.steps(3)
        orc = _hand_over(sim, st)
    else:
        orc = make_oracle(st, 0.4, EPS, 1e-3, 64, flags=1)
    f, a, e2 = sim.coarse_features()
    assert orc.run_stages(1, 4) == 0
    orc.moments()
    srt = orc.sorted()
    assert np.array_equal(srt, sim.sorted())
    oa, of, oe2, ni = orc.coarse(THETA_C, alpha=ALPHA)
    if case.endswith("history"):
        # the relative test is live: it must open cells the extent test takes
        assert ni > orc.coarse(THETA_C, features=False, alpha=0.0)[3]
    n = len(st[0])
    oa, of, oe2 = oa[:n], of[:n], oe2[:n]
    # engine outputs are by body index, oracle rows by sorted position
    ga, gf, ge2 = a[srt], f[srt], e2[srt]
    err = rel_err(ga, oa)
    if case.startswith("massless"):
        # the massive part's pull nearly cancels near its centre of mass (|a|
        # down to 0.05 of the median), where the FP32 sums' rounding, which
        # scales with the cancelling parts, is held to 1e-5 of half the
        # median acceleration instead of the body's own
        on = np.linalg.norm(oa, axis=1)
        err = err * on / np.maximum(on, 0.5 * np.median(on))
    assert err.max() <= 1e-5, (err.max(), int(err.argmax()))
    # features 17 and 18 are logs of integer counts: the counts are the oracle's exactly
    for k in (17, 18):
        cg = np.rint(np.expm1(gf[:, k].astype(np.float64)))
        co = np.rint(np.expm1(of[:, k].astype(np.float64)))
        assert np.array_equal(cg, co), (k, np.argwhere(cg != co)[:5])
    bad = np.abs(gf - of) > 1e-3 + 1e-3 * np.abs(of)
    rows_bad = bad.any(axis=1)
    assert rows_bad.sum() <= max(2, n // 2000), (rows_bad.sum(), np.argwhere(bad)[:10])
    sim.close()


def test_coarse_mode_force_stage_matches_oracle(nb, oracle):
    from nbody_amd import universe as U

    st = U.plummer(20000, seed=12)
    sim = _coarse_sim(st, coarse_alpha=ALPHA)
    sim.steps(2)
    orc = _hand_over(sim, st)
    for s in range(1, 6):
        sim.run_stage(s)
    g = sim.accel()
    assert orc.run_stages(1, 4) == 0
    orc.moments()
    srt = orc.sorted()
    oa = orc.coarse(THETA_C, features=False, alpha=ALPHA)[0][: len(srt)]
    err = rel_err(g[srt], oa)
    assert err.max() <= 1e-5, err.max()
    sim.close()


def test_mlp_bf16_inference_vs_oracle(nb, oracle):
    """Also pins the matrix-instruction lane layout: a wrong A/B/D mapping
    moves outputs by O(1), far outside the BF16 bar."""
    from nbody_amd import universe as U

    st = U.uniform_ball(4096, seed=3)
    sim = _coarse_sim(st, method=nb._lib.NB_MODE_LEARNED)
    blob = _random_model(5)
    sim.set_model(blob)
    assert np.array_equal(sim.get_model(), blob)
    rng = np.random.default_rng(6)
    xmu, xsd = blob[5955:5979], blob[5979:6003]
    ymu, ysd = blob[6003:6006].astype(np.float64), blob[6006:6009].astype(np.float64)
    for rows in (1000, 4096, 77):
        x = (xmu + xsd * rng.normal(size=(rows, 24))).astype(np.float32)
        y = sim.mlp_forward(x)
        xn = np.clip((x - xmu) * (np.float32(1) / xsd), -8, 8).astype(np.float32)
        ref = oracle.mlp_forward(blob[:5955].astype(np.float64), xn) * ysd + ymu
        d = y - ref
        assert np.sqrt((d ** 2).mean()) <= 0.01 * np.sqrt((ref ** 2).mean()), (rows, np.sqrt((d ** 2).mean()))
        assert np.abs(d).max() <= 0.05 * np.abs(ref).max()
    sim.close()


def test_training_gradient_vs_oracle(nb, oracle):
    from nbody_amd import universe as U

    sim = _coarse_sim(U.uniform_ball(2048, seed=3), method=nb._lib.NB_MODE_LEARNED)
    blob = _random_model(7)
    sim.set_model(blob)
    rng = np.random.default_rng(8)
    for n in (512, 1000):
        x = rng.normal(size=(n, 24)).astype(np.float32)
        t = rng.normal(size=(n, 3)).astype(np.float32)
        loss, g = sim.mlp_loss_grad(x, t)
        oloss, og = oracle.mlp_loss_grad(blob[:5955].astype(np.float64), x, t)
        assert abs(loss - oloss) <= 1e-5 * abs(oloss)
        assert np.abs(g - og).max() <= 1e-4 * np.abs(og).max(), np.abs(g - og).max()
    sim.close()


def test_adam_steps(nb, oracle):
    from nbody_amd import universe as U

    sim = _coarse_sim(U.uniform_ball(2048, seed=3), method=nb._lib.NB_MODE_LEARNED)
    blob = _random_model(9)
    sim.set_model(blob)
    rng = np.random.default_rng(10)
    x = rng.normal(size=(256, 24)).astype(np.float32)
    t = rng.normal(size=(256, 3)).astype(np.float32)
    p0 = blob[:5955].astype(np.float64)
    lr = 1e-3
    _, g0 = oracle.mlp_loss_grad(p0, x, t)
    p1 = sim.mlp_adam(x, t, 1, lr)
    big = np.abs(g0) > 1e-5
    # first step from zero moments: -lr * g / (|g| + eps) = -lr sign(g)
    assert np.allclose((p1 - blob[:5955])[big], -lr * np.sign(g0[big]), rtol=1e-3, atol=1e-7)
    # four steps against an FP64 Adam restatement on oracle gradients
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for k in range(1, 5):
        _, g = oracle.mlp_loss_grad(p, x, t)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - lr * (m / (1 - 0.9 ** k)) / (np.sqrt(v / (1 - 0.999 ** k)) + 1e-8)
    p4 = sim.mlp_adam(x, t, 4, lr)
    assert np.abs(p4 - p).max() <= 0.05 * 4 * lr, np.abs(p4 - p).max()
    sim.close()


def test_model_validation(nb):
    from nbody_amd import universe as U

    sim = _coarse_sim(U.uniform_ball(2048, seed=3), method=nb._lib.NB_MODE_LEARNED)
    with pytest.raises(nb.NbError) as e:
        sim.step()                      # learned mode without a model
    assert e.value.code == nb._lib.NB_E_MODEL
    blob = _random_model(1)
    bad = blob.copy()
    bad[6000] = -1.0                    # a negative feature scale
    with pytest.raises(nb.NbError):
        sim.set_model(bad)
    bad = blob.copy()
    bad[10] = np.nan
    with pytest.raises(nb.NbError):
        sim.set_model(bad)
    with pytest.raises(nb.NbError):
        sim.set_model(blob[:-1])
    sim.close()


def test_learned_probe_equals_step_and_fused_equals_unfused(nb):
    from nbody_amd import universe as U

    st = U.plummer(30000, seed=13)
    blob = _random_model(2)
    sims = []
    for flags in (0, nb._lib.NB_FUSED_UPDATE):
        s = _coarse_sim(st, method=nb._lib.NB_MODE_LEARNED, flags=flags)
        s.set_model(blob)
        sims.append(s)
    probe = sims[0].probe_accel()
    for s in sims:
        s.steps(2)
    a0, a1 = sims[0].accel(), sims[1].accel()
    assert np.array_equal(a0, a1)
    x0, x1 = sims[0].state(), sims[1].state()
    for u, w in zip(x0, x1):
        assert np.allclose(u, w, rtol=1e-6, atol=1e-7)
    # the probe ran on the initial positions, as the first step's force did
    sim = _coarse_sim(st, method=nb._lib.NB_MODE_LEARNED)
    sim.set_model(blob)
    sim.steps(1)
    assert np.array_equal(sim.accel(), probe)
    for s in sims + [sim]:
        s.close()


def test_check_fallback_to_plain(nb):
    from nbody_amd import universe as U

    st = U.plummer(20000, seed=14)
    blob = _random_model(3)
    strict = _coarse_sim(st, method=nb._lib.NB_MODE_LEARNED, check_tol=1e-12, check_every=1, check_samples=256)
    strict.set_model(blob)
    strict.steps(2)
    c = strict.counters()
    assert c.fallbacks == 1 and c.last_check_median > 1e-12
    plain = make_sim(st, theta=0.4, softening=EPS, dt=1e-3, vote_width=64)
    plain.steps(2)
    for u, w in zip(strict.state(), plain.state()):
        assert np.allclose(u, w, rtol=1e-5, atol=1e-6)
    assert np.allclose(strict.accel(), plain.accel(), rtol=1e-4, atol=1e-6)
    loose = _coarse_sim(st, method=nb._lib.NB_MODE_LEARNED, check_tol=10.0, check_every=1, check_samples=256)
    loose.set_model(blob)
    loose.steps(2)
    c = loose.counters()
    assert c.fallbacks == 0 and 0.0 < c.last_check_median < 10.0
    for s in (strict, plain, loose):
        s.close()


def test_check_steps_in_one_call(nb):
    """The check's direct sum runs on a second stream beside the step: six
    steps in one call (three check steps, with the host queueing ahead) end
    bit-identical to six single-step calls, and the counters report the
    checks and their device time."""
    from nbody_amd import universe as U

    st = U.plummer(20000, seed=15)
    blob = _random_model(4)
    kw = dict(method=nb._lib.NB_MODE_LEARNED, coarse_alpha=ALPHA, check_tol=10.0, check_every=2,
              check_samples=256, flags=nb._lib.NB_FUSED_UPDATE)
    a = _coarse_sim(st, **kw)
    b = _coarse_sim(st, **kw)
    a.set_model(blob)
    b.set_model(blob)
    a.steps(6)
    for _ in range(6):
        b.steps(1)
    for u, w in zip(a.state(), b.state()):
        assert np.array_equal(u.view(np.uint32), w.view(np.uint32))
    for s in (a, b):
        c = s.counters()
        assert c.checks == 3 and c.check_ms > 0.0 and c.fallbacks == 0
        assert 0.0 < c.last_check_median < 10.0
        s.close()


def test_check_adapt_tightens_alpha(nb):
    """check_adapt: a failed check (tolerance 1e-12 fails every one) makes its
    step plain theta and tightens alpha for the next steps, by at most a factor
    2 each; after three failed checks in a row plain theta holds and checks
    stop. Without check_adapt the first failed check ends the learned force."""
    from nbody_amd import universe as U

    st = U.plummer(20000, seed=16)
    blob = _random_model(5)
    kw = dict(method=nb._lib.NB_MODE_LEARNED, coarse_alpha=ALPHA, check_tol=1e-12, check_every=1,
              check_samples=256)
    a = _coarse_sim(st, check_adapt=1, **kw)
    a.set_model(blob)
    alphas = []
    for _ in range(5):
        a.steps(1)
        c = a.counters()
        alphas.append(c.coarse_alpha)
    c = a.counters()
    assert c.fallbacks == 3 and c.checks == 3, (c.fallbacks, c.checks)
    assert np.isclose(alphas[0], ALPHA * 0.5, rtol=1e-6) and np.isclose(alphas[1], ALPHA * 0.25, rtol=1e-6)
    assert alphas[2] == alphas[1] == alphas[4]  # the third failure ends the learned force instead
    b = _coarse_sim(st, check_adapt=0, **kw)
    b.set_model(blob)
    b.steps(5)
    c = b.counters()
    assert c.fallbacks == 1 and c.checks == 1 and np.isclose(c.coarse_alpha, ALPHA)
    # a loose tolerance leaves alpha alone
    l = _coarse_sim(st, check_adapt=1, **dict(kw, check_tol=10.0))
    l.set_model(blob)
    l.steps(3)
    c = l.counters()
    assert c.fallbacks == 0 and c.checks == 3 and np.isclose(c.coarse_alpha, ALPHA)
    for s in (a, b, l):
        s.close()


@pytest.mark.slow
def test_training_ablation(nb, tmp_path):
    """What the trained network gains over the coarse traversal alone, at the
    benchmark's size (4,194,304 clustered bodies, coarse theta 0.9, alpha
    0.02), on the simulation that produced its training pairs. The gain is a
    few per cent: the extent and relative tests already bound each left-out
    term, and what remains is mostly the higher multipoles of individual
    cells, which no per-body summary sees."""
    from nbody_amd import universe as U

    st = U.gaussian_clusters(4_194_304, seed=4)
    sim = _coarse_sim(st, method=nb._lib.NB_MODE_LEARNED, coarse_alpha=ALPHA)
    rep = sim.train(steps=30, groups_per_step=512, iters=4000, batch=4096, seed=3)
    assert rep["samples"] == 30 * 512 * 64
    # the learnable part of the residual is a few per cent of its variance
    # (normalized units: a network that learns nothing sits at ~1)
    assert rep["loss_last"] < rep["loss_first"], rep
    # the same state, previous accelerations and step counter in both, so the
    # relative test makes the same decisions
    path = str(tmp_path / "state.npz")
    sim.checkpoint(path)
    coarse = _coarse_sim(st, coarse_alpha=ALPHA)
    coarse.restore(path)
    learned = sim.force_error(4096, seed=77)
    alone = coarse.force_error(4096, seed=77)
    print(f"ablation: coarse alone median {alone['median']:.3e} p90 {alone['p90']:.3e}, "
          f"learned median {learned['median']:.3e} p90 {learned['p90']:.3e}, train {rep}")
    assert learned["median"] < 0.99 * alone["median"], (learned, alone)
    assert learned["p90"] < 1.02 * alone["p90"], (learned, alone)
    sim.close()
    coarse.close()
