"""Pins for the CPU oracle, each independent of oracle/bh_oracle.c:

  bbox        numpy min/max closed form (bit-exact) and the reference test's own
              assertion (TestBoundingBox.java:236-251)
  build       hand-derived trees of the reference's test universes
              (EightBodyUniverse, TwoBodyUniverse, BigTreeUniverse); and for
              random inputs a trie built in numpy from each body's own octant
              sequence (a cell exists where two bodies share a path prefix),
              which must give every body the same path and depth as the
              oracle's sequential insertion; insertion-order independence
  summarize   cell counts and FP64 centres of mass recomputed directly from the
              member bodies of every trie prefix
  sort        sorted[] equals the lexicographic order of the octant paths
  force       theta -> 0 equals the FP64 direct sum; two-body closed form; a
              recursive per-body Barnes-Hut written over the numpy trie (vote 1)
  integrate   closed-form kick-drift-kick on given accelerations and the
              first-step rule (no velocity correction at step 0)
  dq table    closed form r^2 / theta^2 / 4^d + eps^2
  direct sum  numpy brute force
"""
import numpy as np
import pytest

from helpers import make_oracle, rel_err

F = np.float32


# ---------------------------------------------------------------- independent trie
def trie_paths(x, y, z, levels=40):
    """Octant of every body at every level, computed per body from the root
    with the reference's FP32 recurrence (buildtree.cl:124-136), vectorized."""
    import pyoracle

    root = pyoracle.bbox(x, y, z)
    n = len(x)
    cx = np.full(n, root[0], F); cy = np.full(n, root[1], F); cz = np.full(n, root[2], F)
    r = np.full(n, root[3], F)
    octs = np.zeros((n, levels), np.int8)
    for L in range(levels):
        bx, by, bz = cx < x, cy < y, cz < z
        octs[:, L] = bx.astype(np.int8) | (by.astype(np.int8) << 1) | (bz.astype(np.int8) << 2)
        rh = (r * F(0.5)).astype(F)
        cx = ((cx - rh).astype(F) + np.where(bx, r, F(0))).astype(F)
        cy = ((cy - rh).astype(F) + np.where(by, r, F(0))).astype(F)
        cz = ((cz - rh).astype(F) + np.where(bz, r, F(0))).astype(F)
        r = rh
    return octs


def trie_depths(octs):
    """Leaf depth of each body: 1 + the longest octant prefix it shares with any
    other body (neighbours in lexicographic order share the longest)."""
    n, L = octs.shape
    order = np.lexsort(octs.T[::-1])
    s = octs[order]
    same = s[1:] == s[:-1]
    lcp = np.where(same.all(axis=1), L, np.argmin(same, axis=1))
    d = np.ones(n, np.int64)
    d[:-1] = np.maximum(d[:-1], lcp + 1)
    d[1:] = np.maximum(d[1:], lcp + 1)
    depth = np.empty(n, np.int64)
    depth[order] = d
    return depth, order


def pack_paths(octs, depth):
    n = len(depth)
    out = np.zeros((n, 3), np.uint64)
    for L in range(octs.shape[1]):
        sel = depth > L
        w, sh = L // 21, 3 * (L % 21)
        out[sel, w] |= octs[sel, L].astype(np.uint64) << np.uint64(sh)
    return out


def _state(n, seed, spread=1.0):
    from nbody_amd import universe as U

    if seed % 2:
        return U.uniform_ball(n, seed=seed)
    return U.gaussian_clusters(n, seed=seed, k=8, box=spread)


# ---------------------------------------------------------------- bbox
@pytest.mark.parametrize("n,seed", [(1, 1), (17, 2), (4096, 3), (32768, 5)])
def test_bbox_closed_form(oracle, n, seed):
    x, y, z = _state(max(n, 2), seed)[:3]
    x, y, z = x[:n], y[:n], z[:n]
    out = oracle.bbox(x, y, z)
    mn = [a.min() for a in (x, y, z)]
    mx = [a.max() for a in (x, y, z)]
    exp = [F(0.5) * (F(mn[k]) + F(mx[k])) for k in range(3)]
    rad = F(0.5) * max(F(mx[0] - mn[0]), F(mx[1] - mn[1]), F(mx[2] - mn[2]))
    assert np.array_equal(out[:3].view(np.uint32), np.array(exp, F).view(np.uint32))
    assert out[3].view(np.uint32) == np.float32(rad).view(np.uint32)


def test_bbox_reference_assertion(oracle):
    """TestBoundingBox: 32768 bodies of RandomCubicUniverseGenerator(32768),
    root centre within 1e-5 of 0.5 (min + max) (TestBoundingBox.java:236-251)."""
    from nbody_amd import universe as U

    st = U.random_cubic(32768, 32768.0, seed=11)
    o = make_oracle(st, 0.5, 0.05, 0.025, 16)
    assert o.run_stages(1, 1) == 0
    root = o.root()
    for k in range(3):
        a = st[k]
        assert abs(F(0.5) * (a.min() + a.max()) - root[k]) < 1e-5


# ---------------------------------------------------------------- build: known answers
def test_eight_body_tree(oracle):
    from nbody_amd import universe as U

    o = make_oracle(U.eight_body(), 0.5, 0.05, 0.025, 16)
    assert o.run_stages(1, 2) == 0
    root = o.root()
    assert np.allclose(root, [-0.05, 0.05, 0.0, 1.1], atol=1e-6)
    p, d = o.paths()
    assert list(d) == [1] * 8
    assert list(p[:, 0]) == [0, 1, 2, 4, 3, 6, 5, 7]
    assert o.scalars()["cells"] == 1


def test_two_body_tree(oracle):
    from nbody_amd import universe as U

    o = make_oracle(U.two_body(), 0.5, 0.05, 0.025, 16)
    assert o.run_stages(1, 2) == 0
    p, d = o.paths()
    assert list(d) == [1, 1] and list(p[:, 0]) == [0, 7]


def test_big_tree(oracle):
    """BigTreeUniverse: root (0,0,0) radius 1e4; bodies 0 and 3 share octant 0
    and split in a cell centred at (-5000,-5000,-5000); body 1 at x = 1e-8 goes
    to octant 1, body 2 to octant 7; body 3 has mass 0."""
    from nbody_amd import universe as U

    o = make_oracle(U.big_tree(), 0.5, 0.05, 0.025, 16)
    assert o.run_stages(1, 2) == 0
    assert np.allclose(o.root(), [0, 0, 0, 1e4])
    p, d = o.paths()
    assert list(d) == [2, 1, 1, 2]
    assert p[1, 0] == 1 and p[2, 0] == 7
    assert p[0, 0] == (0 | 7 << 3) and p[3, 0] == 0
    assert o.scalars()["cells"] == 2
    assert o.run_stages(3, 5) == 0
    assert np.all(np.isfinite(o.accel()))


# ---------------------------------------------------------------- build / sort: trie
@pytest.mark.parametrize("n,seed", [(2, 1), (100, 2), (2048, 3), (5000, 4), (4096, 6)])
def test_paths_match_trie(oracle, n, seed):
    st = _state(n, seed, spread=0.05)
    octs = trie_paths(*st[:3])
    depth, order = trie_depths(octs)
    assert depth.max() < 40
    o = make_oracle(st, 0.5, 0.05, 0.025, 16)
    assert o.run_stages(1, 2) == 0
    p, d = o.paths()
    assert np.array_equal(d, depth)
    assert np.array_equal(p, pack_paths(octs, depth))
    assert o.scalars()["max_depth"] == max(1, depth.max())
    # sort: lexicographic order of the octant paths
    assert o.run_stages(3, 4) == 0
    assert np.array_equal(o.sorted(), order.astype(np.int32))


def test_insertion_order_independence(oracle):
    st = _state(3000, 7)
    perm = np.random.default_rng(0).permutation(3000)
    a = make_oracle(st, 0.5, 0.05, 0.025, 16)
    b = make_oracle(tuple(s[perm] for s in st), 0.5, 0.05, 0.025, 16)
    assert a.run_stages(1, 2) == 0 and b.run_stages(1, 2) == 0
    pa, da = a.paths()
    pb, db = b.paths()
    assert np.array_equal(pa[perm], pb) and np.array_equal(da[perm], db)


# ---------------------------------------------------------------- summarize
def test_summarize_against_direct_sums(oracle):
    from scipy.spatial import cKDTree

    n = 3000
    st = _state(n, 9)
    x, y, z, m = st[0], st[1], st[2], st[6]
    octs = trie_paths(x, y, z)
    depth, _ = trie_depths(octs)
    # every prefix shared by >= 2 bodies is a cell (plus the root)
    cells = {}
    for b in range(n):
        for L in range(depth[b]):  # prefixes of length 0 .. depth-1
            cells.setdefault(octs[b, :L].tobytes() + bytes([L]), []).append(b)
    rows = []
    for members in cells.values():
        idx = np.array(members)
        mm = m[idx].astype(np.float64)
        M = mm.sum()
        c = [np.sum(mm * a[idx].astype(np.float64)) / M for a in (x, y, z)]
        rows.append((len(idx), *c, M))
    rows = np.array(rows)
    o = make_oracle(st, 0.5, 0.05, 0.025, 16)
    assert o.run_stages(1, 3) == 0
    oc = o.cells()
    assert len(oc["count"]) == len(rows)
    assert np.array_equal(np.sort(oc["count"]), np.sort(rows[:, 0].astype(np.int32)))
    assert oc["count"][-1] == n  # root (index NN) holds every body
    got = np.stack([oc["x"], oc["y"], oc["z"]], axis=1).astype(np.float64)
    tree = cKDTree(got)
    dist, j = tree.query(rows[:, 1:4])
    assert np.all(oc["count"][j] == rows[:, 0])
    assert np.all(dist <= 4e-7 * np.maximum(1.0, np.abs(rows[:, 1:4]).max(axis=1)))
    assert np.allclose(oc["m"][j], rows[:, 4], rtol=2e-7)


# ---------------------------------------------------------------- force
def test_direct_fp64_bruteforce(oracle):
    st = _state(500, 11)
    x, y, z, m = (a.astype(np.float64) for a in (st[0], st[1], st[2], st[6]))
    idx = np.arange(0, 500, 7)
    got = oracle.direct_fp64(st[0], st[1], st[2], st[6], 0.05, idx)
    d = np.stack([x[None, :] - x[idx, None], y[None, :] - y[idx, None], z[None, :] - z[idx, None]], -1)
    r2 = (d ** 2).sum(-1) + float(F(0.05) * F(0.05))
    exp = (d * (m[None, :] / r2 ** 1.5)[..., None]).sum(1)
    assert np.allclose(got, exp, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("vote", [1, 16, 64])
def test_theta_zero_is_direct_sum(oracle, vote):
    """THETA -> 0 opens every cell (calculateforce.cl:145): the result is the
    direct sum over all bodies (FP32 separations, so ~1e-7 per term)."""
    n = 2000
    st = _state(n, 13)
    o = make_oracle(st, 1e-6, 0.05, 0.025, vote)
    assert o.run_stages(1, 5) == 0
    a = o.accel()
    ref = oracle.direct_fp64(st[0], st[1], st[2], st[6], 0.05, np.arange(n))
    assert rel_err(a, ref).max() < 2e-6
    assert o.scalars()["interactions"] == n * n


def test_two_body_closed_form(oracle):
    from nbody_amd import universe as U

    st = U.two_body()
    o = make_oracle(st, 0.5, 0.05, 0.025, 16)
    assert o.run_stages(1, 5) == 0
    a = o.accel()
    d = np.array([st[0][1] - st[0][0], st[1][1] - st[1][0], st[2][1] - st[2][0]], np.float64)
    mag = (1 / 8) / (np.dot(d, d) + 0.0025) ** 1.5
    assert np.allclose(a[0], d * mag, rtol=1e-6)
    assert np.allclose(a[1], -d * mag, rtol=1e-6)


def _bh_recursive(x, y, z, m, octs, depth, theta, eps, radius):
    """Per-body Barnes-Hut over the numpy trie (vote width 1): a cell at level
    L (children side radius / 2^L) is accepted when its children's side^2 /
    THETA + eps^2 <= d^2 + eps^2 with FP32 d^2; bodies always interact."""
    n = len(x)
    t2 = F(theta) * F(theta)
    e2 = F(eps) * F(eps)
    dq = [F(F(radius * radius) / t2)]
    for _ in range(63):
        dq.append(F(F(0.25) * dq[-1]))
    dq = [F(v + e2) for v in dq]
    members = {}
    for b in range(n):
        for L in range(depth[b] + 1):
            members.setdefault((L, octs[b, :L].tobytes()), []).append(b)
    com = {}
    for key, mem in members.items():
        idx = np.array(mem)
        mm = m[idx].astype(np.float64)
        M = mm.sum()
        com[key] = (F(np.sum(mm * x[idx]) / M), F(np.sum(mm * y[idx]) / M), F(np.sum(mm * z[idx]) / M), F(M))
    out = np.zeros((n, 3))
    for b in range(n):
        stack = [(0, b"")]
        acc = np.zeros(3)
        while stack:
            L, pre = stack.pop()
            kids = sorted({octs[j, L] for j in members[(L, pre)]})
            for k in kids:
                key = (L + 1, pre + bytes([k]))
                mem = members[key]
                cx_, cy_, cz_, mc = com[key]
                dx, dy, dz = F(cx_ - x[b]), F(cy_ - y[b]), F(cz_ - z[b])
                d2 = np.float64(dx) ** 2 + np.float64(dy) ** 2 + np.float64(dz) ** 2 + np.float64(e2)
                is_body = len(mem) == 1 and depth[mem[0]] == L + 1
                if is_body or d2 >= dq[L]:
                    r = 1 / np.sqrt(d2)
                    acc += np.array([dx, dy, dz], np.float64) * (np.float64(mc) * r ** 3)
                else:
                    stack.append((L + 1, pre + bytes([k])))
        out[b] = acc
    return out


def test_force_matches_recursive_bh(oracle):
    n = 300
    st = _state(n, 15)
    x, y, z, m = st[0], st[1], st[2], st[6]
    octs = trie_paths(x, y, z)
    depth, _ = trie_depths(octs)
    radius = oracle.bbox(x, y, z)[3]
    exp = _bh_recursive(x, y, z, m, octs, depth, 0.5, 0.05, radius)
    o = make_oracle(st, 0.5, 0.05, 0.025, 1)
    assert o.run_stages(1, 5) == 0
    assert rel_err(o.accel(), exp).max() < 1e-5


def test_vote_width_only_opens_more(oracle):
    """A wider vote group can only open more cells: interactions grow and the
    error against the FP64 direct sum does not get worse on the median."""
    n = 8192
    st = _state(n, 17)
    ref = oracle.direct_fp64(st[0], st[1], st[2], st[6], 0.05, np.arange(0, n, 8))
    res = {}
    for w in (1, 16, 64):
        o = make_oracle(st, 0.7, 0.05, 0.025, w)
        assert o.run_stages(1, 5) == 0
        res[w] = (o.scalars()["interactions"], np.median(rel_err(o.accel()[::8], ref)))
    assert res[1][0] < res[16][0] < res[64][0]
    assert res[64][1] <= res[16][1] <= res[1][1]


def test_dq_table_closed_form(oracle):
    r, th, e = F(3.7), 0.4, 0.01
    dq = oracle.dq_table(r, th, e, 20)
    t2 = F(th) * F(th)
    for d in range(20):
        exact = float(r) ** 2 / float(t2) / 4.0 ** d + float(F(e) * F(e))
        assert abs(dq[d] - exact) <= 2 * np.spacing(np.float32(exact))


# ---------------------------------------------------------------- integrate
def test_integrate_closed_form(oracle):
    """integrate.cl:29-43 on given accelerations: v_half = v + a dt/2,
    x = x + v_half dt (one rounding), v = v_half + a dt/2."""
    n = 1000
    g = np.random.default_rng(3)
    st = tuple(g.normal(size=n).astype(F) for _ in range(6)) + (np.full(n, 1e-3, F),)
    a = g.normal(size=(n, 3)).astype(F)
    dt = F(0.025)
    o = make_oracle(st, 0.5, 0.05, float(dt), 16)
    o.set_accel(a[:, 0], a[:, 1], a[:, 2], 0)
    assert o.run_stages(6, 6) == 0
    got = o.state()
    for k in range(3):
        dv = ((a[:, k] * dt) * F(0.5)).astype(F)
        vh = (st[3 + k] + dv).astype(F)
        x = (vh.astype(np.float64) * np.float64(dt) + st[k].astype(np.float64)).astype(F)
        v = (vh + dv).astype(F)
        assert np.max(np.abs(got[k] - x) / np.spacing(np.abs(x))) <= 1
        assert np.array_equal(got[3 + k], v)


def test_first_step_skips_velocity_correction(oracle):
    """step starts at -1 and bbox increments it (GPUBarnesHutNBodySimulation.java:167,
    boundingbox.cl:195): the first force pass sees step 0 and leaves v alone
    (calculateforce.cl:174); later passes add (a_new - a_old) dt / 2."""
    from nbody_amd import universe as U

    st = U.two_body()
    o = make_oracle(st, 0.5, 0.05, 0.025, 16)
    assert o.run_stages(1, 5) == 0
    assert o.scalars()["step"] == 0
    assert np.array_equal(o.state()[3], st[3])
    a0 = o.accel()
    assert o.run_stages(6, 6) == 0
    v_after0 = o.state()[3].copy()
    assert o.run_stages(1, 5) == 0
    a1 = o.accel()
    exp = (v_after0 + ((a1[:, 0] - a0[:, 0]) * F(0.025)) * F(0.5)).astype(F)
    assert np.array_equal(o.state()[3], exp)


def test_two_body_orbit_energy(oracle):
    """Velocity Verlet on a softened circular two-body orbit keeps the energy
    to O(dt^2) over many orbits."""
    m = F(0.5)
    r = 1.0
    v = np.sqrt(0.5 * m / (2 * r))  # each body circles the origin at radius r/2
    x = np.array([-r / 2, r / 2], F); y = np.zeros(2, F); z = np.zeros(2, F)
    vx = np.zeros(2, F); vy = np.array([-v, v], F); vz = np.zeros(2, F)
    st = (x, y, z, vx, vy, vz, np.array([m, m], F))
    o = make_oracle(st, 0.5, 1e-3, 1e-3, 16)

    def energy():
        px, py, pz, ux, uy, uz = o.state()
        k = 0.5 * float(m) * np.sum(ux.astype(float) ** 2 + uy.astype(float) ** 2 + uz.astype(float) ** 2)
        d2 = (px[1] - px[0]) ** 2 + (py[1] - py[0]) ** 2 + (pz[1] - pz[0]) ** 2
        return k - float(m) * float(m) / np.sqrt(float(d2) + 1e-6)

    e0 = energy()
    assert o.step(3000) == 0
    assert abs(energy() - e0) / abs(e0) < 1e-4


def test_nodes_for_reference_sizes(oracle):
    """GPUBarnesHutNBodySimulation.java:219-227 with maxComputeUnits 16."""
    assert oracle.nodes_for(16384) == 32768
    assert oracle.nodes_for(1000) == 16384
    assert oracle.nodes_for(4194304) == 8388608
    assert oracle.nodes_for(8193) == 16400


@pytest.mark.parametrize("vote", [16, 64])
def test_cpu_baseline_fp32_force_is_the_same_traversal(oracle, vote):
    """The CPU baseline's FP32 force pass (oracle/cpu_fp32.c, timed beside the
    GPU) takes the same opening decisions as the oracle (identical interaction
    count) and differs from the FP64-accumulated oracle only by FP32 rounding;
    on a subset of groups it leaves the other rows alone."""
    from nbody_amd import universe as U

    st = U.plummer(20000, seed=41)
    o = oracle.Oracle(len(st[0]), theta=0.5, softening=0.05, dt=1e-3, vote=vote)
    o.set_state(*st)
    assert o.run_stages(1, 5) == 0
    srt = o.sorted()
    ref = o.accel()[srt].astype(np.float64)
    a, ni = o.force_fp32()
    assert ni == o.scalars()["interactions"]
    err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 2e-5, err.max()
    groups = np.array([0, 3, 7], np.int32)
    a2, ni2 = o.force_fp32(groups)
    rows = (groups[:, None] * vote + np.arange(vote)[None, :]).ravel()
    rows = rows[rows < len(st[0])]
    assert np.array_equal(a2[rows], a[rows])
    mask = np.ones(len(st[0]), bool)
    mask[rows] = False
    assert not a2[mask].any() and 0 < ni2 < ni


def test_tight_clump_premise():
    """The parity case tight_run_8219 (helpers.tight_clump) places 27 distinct
    bodies in one cell of level 21: they share their first 21 octants (one
    63-bit key in the key-sorted build, more than its 16-body register path)
    and every other body leaves that cell's path earlier."""
    from helpers import tight_clump

    st = tight_clump()
    pts = np.stack(st[:3], 1)
    assert len(np.unique(pts[-27:], axis=0)) == 27
    orc = make_oracle(st, 0.5, 0.05, 1e-3, 16)
    assert orc.run_stages(1, 2) == 0
    p, d = orc.paths()
    assert np.all(d[-27:] > 21)
    shared = (p[:, 0] == p[-1, 0]) & (d > 21)
    assert shared.sum() == 27 and np.all(shared[-27:])
