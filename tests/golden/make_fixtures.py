"""Writes the golden fixtures of tests/golden/ (run from the repository root):

    python tests/golden/make_fixtures.py

Inputs: the reference's hand-placed test universes (TwoBodyUniverse.java,
EightBodyUniverse.java, BigTreeUniverse.java), a seeded 4096-body ball, and the
reference's two 32,768-body .universe files (universes/*.universe, copied here
unchanged). Expected outputs come from the CPU oracle (oracle/bh_oracle.c, itself
pinned by tests/test_oracle_pins.py) at the reference's shipped constants
(THETA 1.5, EPSILON 0.0025, TIMESTEP 0.025, WARPSIZE 16; calculateforce.cl:11-18):
the root cell, every body's octant path and depth (as a SHA-256 and a depth
histogram, in full for small inputs), the sorted order (SHA-256), the
accelerations of the first force pass and the state after two full steps on up
to 512 sampled bodies.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gpu-nbody_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

THETA = float(np.sqrt(np.float32(1.5)))
SOFTENING = float(np.sqrt(np.float32(0.0025)))
DT = 0.025
VOTE = 16


def cases():
    from nbody_amd import universe as U

    return {
        "two_body": U.two_body(),
        "eight_body": U.eight_body(),
        "big_tree": U.big_tree(),
        "ball_4096": U.uniform_ball(4096, seed=1),
        "sphericaluniverse1": U.read_universe(os.path.join(HERE, "sphericaluniverse1.universe")),
        "montecarlouniverse1": U.read_universe(os.path.join(HERE, "montecarlouniverse1.universe")),
    }


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make(name, st):
    import pyoracle

    n = len(st[0])
    o = pyoracle.Oracle(n, theta=THETA, softening=SOFTENING, dt=DT, vote=VOTE)
    o.set_state(*st)
    # each value is read right after the stage that produces it (summarize
    # overwrites the root's centre with its centre of mass and compacts children)
    assert o.run_stages(1, 1) == 0
    root = o.root()
    assert o.run_stages(2, 2) == 0
    paths, depth = o.paths()
    assert o.run_stages(3, 4) == 0
    srt = o.sorted()
    assert o.run_stages(5, 5) == 0
    acc = o.accel()
    idx = np.arange(n) if n <= 512 else np.sort(np.random.default_rng(7).choice(n, 512, replace=False))
    o2 = pyoracle.Oracle(n, theta=THETA, softening=SOFTENING, dt=DT, vote=VOTE)
    o2.set_state(*st)
    assert o2.step(2) == 0
    x, y, z, vx, vy, vz = o2.state()
    out = dict(root=root, depth_hist=np.bincount(depth, minlength=65).astype(np.int64),
               paths_sha=np.array(sha(paths)), depth_sha=np.array(sha(depth.astype(np.int32))),
               sorted_sha=np.array(sha(srt.astype(np.int32))), idx=idx.astype(np.int32),
               accel=acc[idx].astype(np.float32),
               pos2=np.stack([x, y, z], 1)[idx].astype(np.float32),
               vel2=np.stack([vx, vy, vz], 1)[idx].astype(np.float32))
    if n <= 4096:
        out.update(paths=paths, depth=depth.astype(np.int32), sorted=srt.astype(np.int32))
    if name in ("two_body", "eight_body", "big_tree", "ball_4096"):
        out.update(state=np.stack([np.asarray(a, np.float32) for a in st], 0))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(f"{name}: n={n}, max depth {int(depth.max())}, {os.path.getsize(os.path.join(HERE, name + '.npz'))} bytes")


if __name__ == "__main__":
    for k, v in cases().items():
        make(k, v)
