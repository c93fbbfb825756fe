"""Shared helpers for the parity tests: run the engine and the oracle on the
same seeded input and compare what each stage produced."""
from __future__ import annotations

import numpy as np


def make_sim(state, **kw):
    """Engine handle with `state` uploaded (accelerations 0, step counter -1)."""
    from nbody_amd import HipBarnesHutNBodySimulation, Mode
    from nbody_amd.universe import ArrayUniverse

    n = len(state[0])
    sim = HipBarnesHutNBodySimulation(Mode.DEFAULT, n, ArrayUniverse(state), **kw)
    sim.init(None)
    return sim


def random_model(seed, scale_out=1e-3):
    """A seeded model blob (network parameters, feature and target scalings)
    whose corrections are about scale_out of the coarse acceleration."""
    rng = np.random.default_rng(seed)
    P = np.concatenate([rng.uniform(-1, 1, 24 * 64) / np.sqrt(24), rng.uniform(-1, 1, 64) / np.sqrt(24),
                        rng.uniform(-1, 1, 64 * 64) / 8, rng.uniform(-1, 1, 64) / 8,
                        rng.uniform(-1, 1, 64 * 3) / 8, rng.uniform(-1, 1, 3) / 8])
    xmu = rng.normal(size=24)
    xsd = rng.uniform(0.5, 2.0, 24)
    ymu = rng.normal(size=3) * scale_out
    ysd = rng.uniform(0.5, 2.0, 3) * scale_out
    return np.concatenate([P, xmu, xsd, ymu, ysd]).astype(np.float32)


def make_oracle(state, theta, softening, dt, vote, flags=0):
    import pyoracle

    o = pyoracle.Oracle(len(state[0]), theta=theta, softening=softening, dt=dt, vote=vote, flags=flags)
    o.set_state(*state)
    return o


def rel_err(a, ref):
    """Per-body |a - ref| / |ref| (vector norms)."""
    num = np.linalg.norm(np.asarray(a, np.float64) - np.asarray(ref, np.float64), axis=1)
    den = np.linalg.norm(np.asarray(ref, np.float64), axis=1)
    return num / np.maximum(den, 1e-300)


def cell_multiset(count, x, y, z, m):
    """Live cells (count > 0) as a lexicographically sorted structured array,
    so trees with different cell numbering compare equal bit for bit."""
    keep = count > 0
    rec = np.zeros(int(keep.sum()), dtype=[("count", "<i4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                                            ("m", "<f4")])
    rec["count"], rec["x"], rec["y"], rec["z"], rec["m"] = count[keep], x[keep], y[keep], z[keep], m[keep]
    return np.sort(rec, order=["count", "x", "y", "z", "m"])


def tight_clump(n=8192, seed=9, target=(0.3, -0.2, 0.1)):
    """A uniform ball of n bodies plus 27 bodies on a 3x3x3 grid inside one
    cell of tree level 21 (the cell around `target`, from the root the bounding
    box gives, with the build's FP32 centre recurrence), a quarter of the cell's
    half size apart: the 27 share their first 21 octants and separate a few
    levels deeper."""
    from nbody_amd import universe as U

    f = np.float32
    st = [np.array(a, copy=True) for a in U.uniform_ball(n, seed=seed)]
    c = [f(0.5) * (a.min() + a.max()) for a in st[:3]]
    r = f(0.5) * max(a.max() - a.min() for a in st[:3])
    t = [f(v) for v in target]
    for _ in range(21):
        rh = r * f(0.5)
        c = [(ck - rh) + (r if ck < tk else f(0.0)) for ck, tk in zip(c, t)]
        r = rh
    step = r * f(0.25)
    g = np.array([-1, 0, 1], np.float32) * step
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    pos = [np.float32(ck) + gk.ravel() for ck, gk in zip(c, (gx, gy, gz))]
    m = np.float32(1.0) / np.float32(n + 27)
    out = [np.concatenate([st[k], pos[k]]) for k in range(3)]
    out += [np.concatenate([st[k], np.zeros(27, np.float32)]) for k in range(3, 6)]
    out.append(np.full(n + 27, m, np.float32))
    return tuple(np.ascontiguousarray(a, np.float32) for a in out)
