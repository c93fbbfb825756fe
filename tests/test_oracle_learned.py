"""Pins of the oracle's learned-mode functions (CPU only).

The coarse traversal, its features and the network have no reference
counterpart, so the oracle is checked against what the mathematics fixes:
  moments     the root's second moments equal a brute-force sum over all
              bodies about the global centre of mass (parallel-axis chain)
  extent      the root's extent bounds the distance of every body from the
              centre of mass (brute force), and equals it where every child of
              the cell is a body (closed form)
  quadrupole  a distant cluster seen through one accepted cell: the error of
              monopole + quadrupole falls as R^-3 while the monopole's falls as
              R^-2, also when the softening is comparable to R (which needs the
              trace term of the softened expansion); and the coarse traversal
              converges to the FP64 direct sum much faster than the monopole
  coarse test the relative test is inert before there is a previous
              acceleration (bit-identical with and without it); with alpha -> 0
              it opens down to the bodies (the FP64 direct sum), and a finite
              alpha adds interactions and lowers the error
  features    invariance under scaling all masses (and the previous
              accelerations) by 2 (bit-exact); the identities
              f3^2 + f4^2 = (|a_q| / |a|)^2 = exp(2 f19), f4 >= 0; the group
              sums that vanish: sum |a| (f5 e1 + f6 e2 + f7 e3) = sum (a - mean a)
              and sum (f8 e1 + f9 e2 + f10 e3) = sum (x - centroid) / r_g; and
              the group means taken back out of each body's frame, the same
              vector on every lane and equal to the mean of the per-body parts;
              the counts (17, 18) sum to the traversal's interaction count, are
              the same on every lane of a group, and give the bodies met that
              feature 21's mass split implies at equal masses; exp(2 f20) =
              f0^2 + f1^2 + f2^2; feature 22 against a brute-force group
              radius; feature 23 against the cosine of the monopole part left
              in the returned acceleration
  network     wiring by one-path networks (a transposed or misindexed weight
              moves the output), GELU at known points, and the backward pass
              against central finite differences of the forward pass
"""
import numpy as np
import pytest

F = np.float32


def _tree(oracle, st, theta=0.5, softening=0.01, vote=64):
    n = len(st[0])
    o = oracle.Oracle(n, theta=theta, softening=softening, dt=1e-3, vote=vote, flags=1)
    o.set_state(*st)
    assert o.run_stages(1, 4) == 0
    o.moments()
    return o


def test_root_moments_brute_force(oracle):
    from nbody_amd import universe as U

    st = U.plummer(3000, seed=3)
    o = _tree(oracle, st)
    root = o.n and oracle.nodes_for(o.n)  # the root is node NN
    q = o.cell_moments(root)[:7]
    x, y, z, m = (np.asarray(a, np.float64) for a in (st[0], st[1], st[2], st[6]))
    M = m.sum()
    c = np.array([(m * x).sum(), (m * y).sum(), (m * z).sum()]) / M
    s = np.stack([x - c[0], y - c[1], z - c[2]], 1)
    r2 = (s * s).sum(1)
    Q = np.array([(m * (3 * s[:, 0] ** 2 - r2)).sum(), (m * (3 * s[:, 1] ** 2 - r2)).sum(),
                  (m * (3 * s[:, 2] ** 2 - r2)).sum(), (m * 3 * s[:, 0] * s[:, 1]).sum(),
                  (m * 3 * s[:, 0] * s[:, 2]).sum(), (m * 3 * s[:, 1] * s[:, 2]).sum(), (m * r2).sum()])
    scale = np.abs(Q).max()
    assert np.allclose(q, Q, rtol=0, atol=1e-10 * scale), (q, Q)
    assert abs(q[0] + q[1] + q[2]) < 1e-10 * scale  # traceless


def _cluster_and_probe(R, eps, seed=0):
    """A 40-body cluster (sigma 0.08) centred in one depth-6 cell of a root
    fixed to [-50, 50]^3 by two light corner bodies, and one light probe body
    at distance R from the cluster's centre. The probe sits outside that cell,
    so it meets the whole cluster as one accepted cell (or as an ancestor that
    holds nothing else)."""
    rng = np.random.default_rng(seed)
    k = 40
    c0 = -50.0 + (50.0 / 64.0) * 41          # centre of a depth-6 cell
    pts = c0 + rng.normal(size=(k, 3)) * 0.08
    u = np.array([0.48, 0.6, 0.64])
    u /= np.linalg.norm(u)
    probe = c0 + u * R
    corners = np.array([[-50.0, -50.0, -50.0], [50.0, 50.0, 50.0]])
    P = np.vstack([pts, probe[None], corners]).astype(np.float32)
    m = np.concatenate([rng.uniform(0.5, 1.5, k), [1e-9, 1e-9, 1e-9]]).astype(np.float32)
    z = np.zeros(len(P), np.float32)
    return (P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy(), z, z, z, m), k


def _probe_error(oracle, R, eps):
    st, k = _cluster_and_probe(R, eps)
    n = len(st[0])
    o = _tree(oracle, st, theta=1.2, softening=eps, vote=1)
    srt = o.sorted()
    acc, _, _, _ = o.coarse(1.2, features=False, geometric=True)   # monopole + quadrupole (vote 1: per body)
    pos = int(np.where(srt == k)[0][0])            # the probe's sorted position
    a_quad = acc[pos]
    # the monopole alone: the plain force stage of the oracle with the same table
    o.run_stages(5, 5)
    a_mono = o.accel()[k].astype(np.float64)
    # exact softened sum over the cluster (the anchor's pull is the same in all three)
    x, y, zz, m = (np.asarray(a, np.float64) for a in (st[0], st[1], st[2], st[6]))
    exact = oracle.direct_fp64(st[0], st[1], st[2], st[6], eps, np.array([k]))[0]
    e_q = np.linalg.norm(a_quad - exact) / np.linalg.norm(exact)
    e_m = np.linalg.norm(a_mono - exact) / np.linalg.norm(exact)
    return e_m, e_q


@pytest.mark.parametrize("eps", [1e-3, 0.4])
def test_quadrupole_convergence_order(oracle, eps):
    """Monopole error ~ R^-2, monopole + softened quadrupole ~ R^-3."""
    em1, eq1 = _probe_error(oracle, 1.5, eps)
    em2, eq2 = _probe_error(oracle, 3.0, eps)
    assert eq1 < 0.5 * em1 and eq2 < 0.3 * em2
    assert 3.0 < em1 / em2 < 5.5, (em1, em2)       # second order
    assert eq1 / eq2 > 6.0, (eq1, eq2)             # third order


def test_coarse_traversal_converges_faster_than_monopole(oracle):
    from nbody_amd import universe as U

    st = U.gaussian_clusters(8192, seed=4)
    o = _tree(oracle, st, theta=0.3)
    srt = o.sorted()
    ref = oracle.direct_fp64(st[0], st[1], st[2], st[6], 0.01, srt)
    med = lambda a: np.median(np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1))
    acc, _, _, _ = o.coarse(0.3, features=False, geometric=True)
    o.run_stages(5, 5)
    mono = o.accel()[srt].astype(np.float64)
    assert med(acc) < 0.25 * med(mono), (med(acc), med(mono))


def test_extent_bounds_bodies_and_closed_forms(oracle):
    from nbody_amd import universe as U

    # brute force: the root's extent bounds every body's distance from the
    # centre of mass, and is within the triangle-inequality slack of it
    st = U.plummer(3000, seed=3)
    o = _tree(oracle, st)
    root = oracle.nodes_for(o.n)
    b = o.cell_moments(root)[7]
    x, y, z, m = (np.asarray(a, np.float64) for a in (st[0], st[1], st[2], st[6]))
    c = np.array([(m * x).sum(), (m * y).sum(), (m * z).sum()]) / m.sum()
    rmax = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2).max()
    assert rmax * (1 - 1e-12) <= b <= 3.0 * rmax, (b, rmax)
    # closed form: one body in each octant of the root (every child is a body)
    pts = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64)
    pts = pts * np.array([0.3, 0.5, 0.7]) + np.array([0.01, -0.02, 0.03]) * np.arange(8)[:, None]
    mm = np.linspace(1.0, 2.0, 8)
    st8 = tuple(np.ascontiguousarray(a, np.float32) for a in (pts[:, 0], pts[:, 1], pts[:, 2], np.zeros(8),
                                                              np.zeros(8), np.zeros(8), mm))
    o8 = _tree(oracle, st8)
    p32 = np.stack(st8[:3], 1).astype(np.float64)
    m32 = st8[6].astype(np.float64)
    cell = oracle.nodes_for(8)
    # the root's centre of mass as the oracle sums it (pairwise over the slots)
    com = o8.cells()
    cm = np.array([com["x"][-1], com["y"][-1], com["z"][-1]], np.float64)
    exact = np.sqrt(((p32 - (m32[:, None] * p32).sum(0) / m32.sum()) ** 2).sum(1)).max()
    assert abs(o8.cell_moments(cell)[7] - exact) <= 1e-12 * exact, (o8.cell_moments(cell)[7], exact, cm)


def test_relative_test_inert_without_history_and_exact_in_the_limit(oracle):
    from nbody_amd import universe as U

    n = 4096
    st = U.gaussian_clusters(n, seed=4)
    o = _tree(oracle, st)
    srt = o.sorted()
    # step 0 after the tree stages: no previous acceleration, alpha changes nothing
    a0, f0, _, n0 = o.coarse(0.9, alpha=0.0)
    a1, f1, _, n1 = o.coarse(0.9, alpha=0.02)
    assert np.array_equal(a0, a1) and np.array_equal(f0, f1) and n0 == n1
    # with previous accelerations and alpha = 1e-18, a cell is taken only where
    # its left-out term m b^3 / s^5 is below 1e-18 |a|: the FP64 direct sum
    ref = oracle.direct_fp64(st[0], st[1], st[2], st[6], 0.01, np.arange(n))
    o2 = oracle.Oracle(n, theta=0.5, softening=0.01, dt=1e-3, vote=64, flags=1)
    o2.set_state(*st)
    o2.set_accel(*(np.ascontiguousarray(ref[:, j], np.float32) for j in range(3)), 1)
    assert o2.run_stages(1, 4) == 0
    a2, _, _, n2 = o2.coarse(0.9, features=False, alpha=1e-18)
    assert n2 > 0.99 * n * n
    e = np.linalg.norm(a2[:n] - ref[srt], axis=1) / np.linalg.norm(ref[srt], axis=1)
    assert e.max() < 1e-6, e.max()            # FP32 offsets d, as in the force stage
    # alpha = 0.02 opens more than the extent test alone, for a smaller error
    a3, _, _, n3 = o2.coarse(0.9, features=False, alpha=0.02)
    a4, _, _, n4 = o2.coarse(0.9, features=False, alpha=0.0)
    med = lambda a: np.median(np.linalg.norm(a[:n] - ref[srt], axis=1) / np.linalg.norm(ref[srt], axis=1))
    assert n3 > n4 and med(a3) < med(a4)


def test_features_mass_scaling_invariance_and_identities(oracle):
    from nbody_amd import universe as U

    st = list(U.plummer(4096, seed=5))
    n = len(st[0])

    def run(state, scale):
        o = oracle.Oracle(n, theta=0.5, softening=0.01, dt=1e-3, vote=64, flags=1)
        o.set_state(*state)
        # previous accelerations: the geometric traversal of a first tree
        assert o.run_stages(1, 4) == 0
        ap, _, _, _ = o.coarse(0.5, features=False, geometric=True)
        srt = o.sorted()
        prev = np.zeros((n, 3), np.float32)
        prev[srt] = ap[:n]
        o.set_accel(*(np.ascontiguousarray(prev[:, j]) for j in range(3)), 1)
        assert o.run_stages(1, 4) == 0
        return o.coarse(0.9, alpha=0.02)

    a1, f1, e21, n1 = run(st, 1)
    st2 = st.copy()
    st2[6] = (st[6] * F(2)).astype(np.float32)
    a2, f2, e22, n2 = run(st2, 2)
    assert n1 == n2
    assert np.array_equal(a2, 2 * a1)
    assert np.array_equal(f1, f2)
    assert np.array_equal(e21, e22)
    # frame identities
    assert (f1[:, 4] >= 0).all()
    lhs = f1[:, 3].astype(np.float64) ** 2 + f1[:, 4].astype(np.float64) ** 2
    rhs = np.exp(2 * f1[:, 19].astype(np.float64))
    assert np.allclose(lhs, rhs, rtol=1e-4, atol=0)
    # e2 is a unit vector orthogonal to the coarse acceleration
    assert np.allclose(np.linalg.norm(e21, axis=1), 1.0, atol=1e-12)
    assert np.abs((e21 * a1).sum(1)).max() <= 1e-12 * np.linalg.norm(a1, axis=1).max()
    # group sums that vanish: deviations from the group's mean acceleration and
    # offsets from its centroid, taken back out of the frame
    A = np.linalg.norm(a1, axis=1, keepdims=True)
    e1 = a1 / A
    e3 = np.cross(e1, e21)
    back = lambda f: f[:, 0:1] * e1 + f[:, 1:2] * e21 + f[:, 2:3] * e3
    dev = (back(f1[:, 5:8].astype(np.float64)) * A).reshape(-1, 64, 3).sum(1)
    assert np.abs(dev).max() <= 1e-5 * A.max(), np.abs(dev).max()
    off = back(f1[:, 8:11].astype(np.float64)).reshape(-1, 64, 3).sum(1)
    assert np.abs(off).max() <= 1e-4, np.abs(off).max()
    # a_b from features 0-2, a_q from 3-4; the group means from 11-13 and 14-16
    ab = (back(f1[:, 0:3].astype(np.float64)) * A).reshape(-1, 64, 3)
    aq = ((f1[:, 3:4] * e1 + f1[:, 4:5] * e21) * A).reshape(-1, 64, 3)
    for part, cols in ((ab, slice(11, 14)), (aq, slice(14, 17))):
        gm = (back(f1[:, cols].astype(np.float64)) * A).reshape(-1, 64, 3)
        scale = np.linalg.norm(part.mean(1), axis=1).max() + 1e-300
        assert np.abs(gm - part.mean(1, keepdims=True)).max() <= 1e-5 * scale


def test_feature_counts_radius_and_cosine(oracle):
    """Features 17, 18, 20, 22 and 23 against what fixes them from outside
    the feature code: the interaction count the traversal returns, the mass
    split of feature 21 (equal masses: bodies met = N (1 - f21)), a
    brute-force group radius, the body part (features 0-2), and the monopole
    part left when the body and quadrupole parts are taken out of the
    returned coarse acceleration."""
    from nbody_amd import universe as U

    st = U.plummer(4096, seed=6)  # equal masses 1/N
    n = len(st[0])
    o = oracle.Oracle(n, theta=0.5, softening=0.01, dt=1e-3, vote=64, flags=1)
    o.set_state(*st)
    assert o.run_stages(1, 4) == 0
    ap, _, _, _ = o.coarse(0.5, features=False, geometric=True)
    srt = o.sorted()
    prev = np.zeros((n, 3), np.float32)
    prev[srt] = ap[:n]
    o.set_accel(*(np.ascontiguousarray(prev[:, j]) for j in range(3)), 1)
    assert o.run_stages(1, 4) == 0
    srt = o.sorted()
    a, f, e2, ninter = o.coarse(0.9, alpha=0.02)
    f = f.astype(np.float64)
    # 17, 18: integer counts, the same on every lane of a vote group (the
    # group takes a cell or opens it as a whole), summing to the interactions
    nc = np.expm1(f[:, 17])
    nb = np.expm1(f[:, 18])
    assert np.abs(nc - np.rint(nc)).max() < 1e-2 and np.abs(nb - np.rint(nb)).max() < 1e-2
    nc, nb = np.rint(nc).astype(np.int64), np.rint(nb).astype(np.int64)
    assert (nc.reshape(-1, 64) == nc.reshape(-1, 64)[:, :1]).all()
    assert (nb.reshape(-1, 64) == nb.reshape(-1, 64)[:, :1]).all()
    assert int((nc + nb).sum()) == ninter
    assert np.abs(nb - n * (1.0 - f[:, 21])).max() < 0.05
    assert (nc > 0).all() and (nb > 0).all()
    # 20: |a_b| / |a| from the frame components of a_b
    assert np.allclose(np.exp(2 * f[:, 20]), (f[:, 0:3] ** 2).sum(1), rtol=1e-4, atol=0)
    # 22: the group's radius about its centroid over the root's radius
    P = np.stack(st[:3], 1).astype(np.float64)[srt].reshape(-1, 64, 3)
    cen = P.mean(1, keepdims=True)
    rg = np.sqrt(((P - cen) ** 2).sum(2)).max(1)
    radius = float(o.root()[3])
    assert np.allclose(np.exp(f[:, 22]).reshape(-1, 64), (rg / radius)[:, None], rtol=1e-5, atol=0)
    # 23: cosine between the cell monopole part and the body part
    A = np.linalg.norm(a, axis=1, keepdims=True)
    e1 = a / A
    e3 = np.cross(e1, e2)
    ab = (f[:, 0:1] * e1 + f[:, 1:2] * e2 + f[:, 2:3] * e3) * A
    aq = (f[:, 3:4] * e1 + f[:, 4:5] * e2) * A
    mono = a - ab - aq
    cos = (mono * ab).sum(1) / (np.linalg.norm(mono, axis=1) * np.linalg.norm(ab, axis=1))
    assert np.abs(cos - f[:, 23]).max() < 1e-4, np.abs(cos - f[:, 23]).max()
    assert f[:, 23].min() >= -1.0 and f[:, 23].max() <= 1.0


# ---- network ---------------------------------------------------------------
def _gelu_tanh(x):
    return 0.5 * x * (1 + np.tanh(np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3)))


def _offsets():
    W1 = 0
    b1 = W1 + 24 * 64
    W2 = b1 + 64
    b2 = W2 + 64 * 64
    W3 = b2 + 64
    b3 = W3 + 64 * 3
    return W1, b1, W2, b2, W3, b3


def test_mlp_one_path_wiring(oracle):
    W1, b1, W2, b2, W3, b3 = _offsets()
    P = np.zeros(oracle.MLP_PARAMS)
    P[W1 + 5 * 64 + 7] = 1.0      # input 5 -> hidden-1 unit 7
    P[W2 + 7 * 64 + 3] = 1.0      # hidden-1 7 -> hidden-2 3
    P[W3 + 3 * 3 + 2] = 1.0       # hidden-2 3 -> output 2
    P[b3 + 0] = 0.25
    x = np.zeros((4, 24), np.float32)
    x[:, 5] = [-2.0, 0.0, 1.0, 3.0]
    x[:, 6] = 9.0                 # an unconnected input changes nothing
    y = oracle.mlp_forward(P, x)
    assert np.allclose(y[:, 0], 0.25) and np.allclose(y[:, 1], 0.0)
    assert np.allclose(y[:, 2], _gelu_tanh(_gelu_tanh(x[:, 5].astype(np.float64))), rtol=1e-12, atol=1e-15)


def test_gelu_known_points(oracle):
    W1, b1, W2, b2, W3, b3 = _offsets()
    P = np.zeros(oracle.MLP_PARAMS)
    P[W1 + 0] = 1.0
    P[W2 + 0] = 1.0
    P[W3 + 0] = 1.0
    # with b2 pushing the second layer far positive, gelu(z) = z there
    P[b2 + 0] = 50.0
    x = np.zeros((3, 24), np.float32)
    x[:, 0] = [0.0, 1.0, -30.0]
    y = oracle.mlp_forward(P, x)[:, 0] - 50.0
    assert abs(y[0]) < 1e-12                      # gelu(0) = 0
    assert abs(y[1] - 0.8411919906082768) < 1e-9  # tanh-form GELU(1)
    assert abs(y[2]) < 1e-12                      # gelu(-30) = 0


def test_mlp_gradient_finite_differences(oracle):
    rng = np.random.default_rng(11)
    P = rng.normal(size=oracle.MLP_PARAMS) * 0.3
    x = rng.normal(size=(16, 24)).astype(np.float32)
    t = rng.normal(size=(16, 3)).astype(np.float32)
    loss, g = oracle.mlp_loss_grad(P, x, t)
    assert loss > 0
    idx = np.concatenate([rng.choice(oracle.MLP_PARAMS, 60, replace=False), [5952, 5953, 5954, 1536, 5696]])
    h = 1e-6
    for i in idx:
        Pp = P.copy(); Pp[i] += h
        Pm = P.copy(); Pm[i] -= h
        fd = (oracle.mlp_loss_grad(Pp, x, t)[0] - oracle.mlp_loss_grad(Pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
