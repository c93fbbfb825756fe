"""Golden fixtures (tests/golden/, written by tests/golden/make_fixtures.py).

CPU: the .universe reader against the reference's own data files (body count,
the unit-sphere radius and equal masses they were generated with, and a
byte-identical rewrite through write_universe), and the oracle reproducing
every stored expected value exactly (a change in the oracle shows up here).

GPU: the engine at the reference's shipped constants on the same inputs:
root cell bit-exact, every body's octant path, depth and sorted position
identical (SHA-256 of the full arrays), accelerations of the first force pass
within 1e-5 relative, state after two full steps within 1e-6.
"""
import hashlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = ["two_body", "eight_body", "big_tree", "ball_4096", "sphericaluniverse1", "montecarlouniverse1"]
THETA = float(np.sqrt(np.float32(1.5)))
SOFTENING = float(np.sqrt(np.float32(0.0025)))
DT, VOTE = 0.025, 16


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _inputs(name):
    from nbody_amd import universe as U

    if name.endswith("universe1"):
        return U.read_universe(os.path.join(GOLD, name + ".universe"))
    g = np.load(os.path.join(GOLD, name + ".npz"))
    return tuple(np.ascontiguousarray(a) for a in g["state"])


@pytest.mark.parametrize("name", ["sphericaluniverse1", "montecarlouniverse1"])
def test_reference_universe_files(name, tmp_path):
    from nbody_amd import universe as U

    path = os.path.join(GOLD, name + ".universe")
    st = U.read_universe(path)
    assert len(st) == 7 and all(len(a) == 32768 for a in st)
    r = np.sqrt(st[0].astype(np.float64) ** 2 + st[1] ** 2 + st[2] ** 2)
    assert np.allclose(r, 1.0, atol=1e-6)                 # points on the unit sphere
    assert np.all(st[3] == 0) and np.all(st[4] == 0) and np.all(st[5] == 0)
    assert np.all(st[6] == np.float32(1.0 / 32768))
    out = tmp_path / "copy.universe"
    U.write_universe(str(out), st)
    assert out.read_bytes() == open(path, "rb").read()   # the reference's own layout, byte for byte


def test_universe_reader_older_layout(tmp_path):
    """The older layout: N as a boxed java.lang.Integer, arrays longer than N."""
    from nbody_amd import universe as U

    n, length = 3, 5
    b = bytearray(b"\xac\xed\x00\x05")
    b += bytes([0x73, 0x72]) + (17).to_bytes(2, "big") + b"java.lang.Integer" + bytes.fromhex("12e2a0a4f7818738")
    b += bytes([0x02]) + (1).to_bytes(2, "big") + b"I" + (5).to_bytes(2, "big") + b"value" + bytes([0x78])
    b += bytes([0x72]) + (16).to_bytes(2, "big") + b"java.lang.Number" + bytes.fromhex("86ac951d0b94e08b")
    b += bytes([0x02, 0x00, 0x00, 0x78, 0x70]) + n.to_bytes(4, "big")
    vals = np.arange(7 * length, dtype=np.float32).reshape(7, length)
    for k in range(7):
        b += bytes([0x75])
        if k == 0:
            b += bytes([0x72]) + (2).to_bytes(2, "big") + b"[F" + bytes.fromhex("0b9c818922e00c42") + bytes(
                [0x02, 0x00, 0x00, 0x78, 0x70])
        else:
            b += bytes([0x71]) + (0x7E0002).to_bytes(4, "big")
        b += length.to_bytes(4, "big") + vals[k].astype(">f4").tobytes()
    p = tmp_path / "old.universe"
    p.write_bytes(bytes(b))
    st = U.read_universe(str(p))
    assert all(np.array_equal(st[k], vals[k, :n]) for k in range(7))


@pytest.mark.parametrize("name", CASES)
def test_fixture_roots_closed_form(name):
    """The stored root is 0.5 (min + max) per axis and half the largest extent
    (boundingbox.cl:171-189), computed here in FP32 from the inputs."""
    st = _inputs(name)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    mn = np.array([a.min() for a in st[:3]], np.float32)
    mx = np.array([a.max() for a in st[:3]], np.float32)
    centre = np.float32(0.5) * (mn + mx)
    assert np.array_equal(g["root"][:3], centre)
    assert g["root"][3] == np.float32(0.5) * np.max(mx - mn)


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixtures(oracle, name):
    st = _inputs(name)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    n = len(st[0])
    o = oracle.Oracle(n, theta=THETA, softening=SOFTENING, dt=DT, vote=VOTE)
    o.set_state(*st)
    assert o.run_stages(1, 1) == 0
    assert np.array_equal(o.root().view(np.uint32), g["root"].view(np.uint32))
    assert o.run_stages(2, 2) == 0
    paths, depth = o.paths()
    assert _sha(paths) == str(g["paths_sha"]) and _sha(depth.astype(np.int32)) == str(g["depth_sha"])
    assert np.array_equal(np.bincount(depth, minlength=65), g["depth_hist"])
    assert o.run_stages(3, 4) == 0
    assert _sha(o.sorted().astype(np.int32)) == str(g["sorted_sha"])
    assert o.run_stages(5, 5) == 0
    assert np.array_equal(o.accel()[g["idx"]], g["accel"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_engine_matches_fixtures(nb, name):
    from helpers import make_sim, rel_err

    st = _inputs(name)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    kw = dict(theta=THETA, softening=SOFTENING, dt=DT, vote_width=VOTE)
    sim = make_sim(st, **kw)
    sim.run_stage(1)
    assert np.array_equal(sim.root().view(np.uint32), g["root"].view(np.uint32))
    sim.run_stage(2)
    paths, depth = sim.paths()
    assert _sha(paths) == str(g["paths_sha"]), "octant paths differ"
    assert _sha(depth.astype(np.int32)) == str(g["depth_sha"]), "depths differ"
    sim.run_stage(3)
    sim.run_stage(4)
    assert _sha(sim.sorted().astype(np.int32)) == str(g["sorted_sha"])
    sim.run_stage(5)
    idx = g["idx"]
    a = sim.accel()[idx]
    err = rel_err(a, g["accel"])
    assert err.max() <= 1e-5, err.max()
    sim.close()
    sim = make_sim(st, **kw)
    sim.steps(2)
    x, y, z, vx, vy, vz = sim.state()
    pos = np.stack([x, y, z], 1)[idx]
    vel = np.stack([vx, vy, vz], 1)[idx]
    for got, ref in ((pos, g["pos2"]), (vel, g["vel2"])):
        tol = 1e-6 * np.abs(ref).max() + 1e-6 * np.abs(ref)
        assert np.all(np.abs(got - ref) <= tol), np.abs(got - ref).max()
    sim.close()
