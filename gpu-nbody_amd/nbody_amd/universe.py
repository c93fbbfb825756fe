"""Seeded initial conditions.

Two families:
  * the benchmark universes (uniform rotating ball, Plummer sphere, two
    Plummer spheres, Gaussian clusters, halo + disk), generated with numpy's
    Philox counter-based generator from a fixed seed;
  * seeded ports of the reference's generators (RandomCubicUniverseGenerator,
    SphericalUniverseGenerator, MonteCarloSphericalUniverseGenerator,
    LonLatSphericalUniverseGenerator, PlummerUniverseGenerator,
    RotatingDiskGalaxyGenerator and the hand-placed test universes), with the
    reference's unseeded Math.random() replaced by that generator.

Every generator returns float32 arrays (x, y, z, vx, vy, vz, m) and also
implements UniverseGenerator.generate (UniverseGenerator.java:22) so it can be
handed to HipBarnesHutNBodySimulation.
"""
from __future__ import annotations

import math
from typing import Callable, Tuple

import numpy as np

State = Tuple[np.ndarray, ...]


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def _f32(*arrs):
    return tuple(np.ascontiguousarray(a, dtype=np.float32) for a in arrs)


def _isotropic(g: np.random.Generator, n: int) -> np.ndarray:
    v = g.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


# ---- benchmark universes ------------------------------------------------------------

def uniform_ball(n: int, seed: int = 1, omega: float = 0.5, sigma_v: float = 0.01) -> State:
    """Uniform in the unit ball, equal masses 1/N, solid-body rotation about z
    (v = omega z x r) plus N(0, sigma_v^2) noise per component."""
    g = rng(seed)
    r = g.random(n) ** (1.0 / 3.0)
    p = _isotropic(g, n) * r[:, None]
    v = np.stack([-omega * p[:, 1], omega * p[:, 0], np.zeros(n)], axis=1)
    v += g.normal(scale=sigma_v, size=(n, 3))
    return _f32(p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], np.full(n, 1.0 / n))


def _plummer_shells(g: np.random.Generator, n: int, a: float, mass: float):
    """Plummer sphere (scale a, total mass `mass`, G = 1) in isotropic
    equilibrium: radii by inverting the cumulative mass with X ~ U(0, 0.999)
    (the truncation of PlummerUniverseGenerator.java:12) and speeds q * v_esc
    with q drawn from q^2 (1 - q^2)^(7/2) by rejection (java:26-30)."""
    X = g.random(n) * 0.999
    X = np.maximum(X, 1e-300)
    r = a / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    p = _isotropic(g, n) * r[:, None]
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x = g.random(todo.size)
        y = g.random(todo.size) * 0.1
        ok = y <= x * x * (1.0 - x * x) ** 3.5
        q[todo[ok]] = x[ok]
        todo = todo[~ok]
    speed = q * np.sqrt(2.0 * mass / a) * (1.0 + (r / a) ** 2) ** -0.25
    v = _isotropic(g, n) * speed[:, None]
    return p, v


def plummer(n: int, seed: int = 2, a: float = 1.0, total_mass: float = 1.0) -> State:
    """Plummer sphere, scale radius a, total mass M, isotropic equilibrium velocities."""
    g = rng(seed)
    p, v = _plummer_shells(g, n, a, total_mass)
    return _f32(p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], np.full(n, total_mass / n))


def two_plummer(n: int, seed: int = 3, offset: float = 5.0, v_rel: float = 0.3) -> State:
    """Two Plummer spheres of n/2 bodies (a = 1, mass 1/2 each) centred at
    x = -offset and x = +offset, moving with x-velocity +v_rel and -v_rel."""
    g = rng(seed)
    h = n // 2
    pa, va = _plummer_shells(g, h, 1.0, 0.5)
    pb, vb = _plummer_shells(g, n - h, 1.0, 0.5)
    pa[:, 0] -= offset
    pb[:, 0] += offset
    va[:, 0] += v_rel
    vb[:, 0] -= v_rel
    p = np.concatenate([pa, pb])
    v = np.concatenate([va, vb])
    return _f32(p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], np.full(n, 1.0 / n))


def gaussian_clusters(n: int, seed: int = 4, k: int = 64, box: float = 10.0) -> State:
    """k Gaussian clusters: centres uniform in [-box, box]^3, sigma_k
    log-uniform in [0.01, 0.2], lognormal(0, 1) mass weights apportioned to
    body counts by largest remainder, equal body masses 1/N, isotropic Gaussian
    velocities with per-component variance M_k / (6 sigma_k)."""
    g = rng(seed)
    centres = g.uniform(-box, box, size=(k, 3))
    sig = np.exp(g.uniform(math.log(0.01), math.log(0.2), size=k))
    w = g.lognormal(0.0, 1.0, size=k)
    w /= w.sum()
    raw = w * n
    cnt = np.floor(raw).astype(np.int64)
    rem = n - int(cnt.sum())
    cnt[np.argsort(-(raw - cnt), kind="stable")[:rem]] += 1
    lab = np.repeat(np.arange(k), cnt)
    p = centres[lab] + g.normal(size=(n, 3)) * sig[lab, None]
    mk = cnt / n
    vs = np.sqrt(mk / (6.0 * sig))
    v = g.normal(size=(n, 3)) * vs[lab, None]
    return _f32(p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], np.full(n, 1.0 / n))


def halo_disk(n: int, seed: int = 5, n_disk: int | None = None, a: float = 1.0,
              r_d: float = 3.0, h_z: float = 0.3) -> State:
    """Plummer halo (3/4 of the bodies, scale a) and an exponential disk
    (1/4, scale length r_d, scale height h_z) on circular orbits computed from
    the mass enclosed in a sphere of the cylindrical radius. Equal masses 1/N."""
    g = rng(seed)
    nd = n // 4 if n_disk is None else n_disk
    nh = n - nd
    mh, md = nh / n, nd / n
    ph, vh = _plummer_shells(g, nh, a, mh)
    R = r_d * g.gamma(2.0, 1.0, size=nd)
    phi = g.uniform(0.0, 2.0 * math.pi, size=nd)
    z = g.laplace(0.0, h_z, size=nd)
    pd = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    menc = mh * R ** 3 / (R * R + a * a) ** 1.5 + md * (1.0 - (1.0 + R / r_d) * np.exp(-R / r_d))
    vc = np.sqrt(menc / np.maximum(R, 1e-12))
    vd = np.stack([-vc * np.sin(phi), vc * np.cos(phi), np.zeros(nd)], axis=1)
    p = np.concatenate([ph, pd])
    v = np.concatenate([vh, vd])
    return _f32(p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], np.full(n, 1.0 / n))


CONFIGS: dict = {
    "uniform_ball": uniform_ball,
    "plummer": plummer,
    "two_plummer": two_plummer,
    "gaussian_clusters": gaussian_clusters,
    "halo_disk": halo_disk,
}


# ---- seeded ports of the reference generators ------------------------------------

def random_cubic(n: int, range_: float, seed: int = 0) -> State:
    """RandomCubicUniverseGenerator.java:12-19: uniform in a cube of side
    `range_` centred at the origin, masses 1/N, zero velocity."""
    g = rng(seed)
    u = g.random((3, n))
    x, y, z = ((u - 0.5) * range_).astype(np.float32)
    zero = np.zeros(n, np.float32)
    return _f32(x, y, z, zero, zero, zero, np.full(n, np.float32(1.0) / np.float32(n)))


def spherical(n: int, seed: int = 0) -> State:
    """SphericalUniverseGenerator.java:6-20: uniform on the unit sphere."""
    g = rng(seed)
    omega = g.random(n) * 2 * math.pi
    u = g.random(n) * 2 - 1
    s = np.sqrt(1 - u * u)
    zero = np.zeros(n)
    return _f32(s * np.cos(omega), s * np.sin(omega), u, zero, zero, zero, np.full(n, 1.0 / n))


def monte_carlo_spherical(n: int, seed: int = 0) -> State:
    """MonteCarloSphericalUniverseGenerator.java:6-27: Fibonacci sphere with a
    random phase."""
    g = rng(seed)
    rand = g.random() * n
    o = 2.0 / n
    inc = math.pi * (3.0 - math.sqrt(5))
    i = np.arange(n, dtype=np.float64)
    y = ((i * o) - 1) + (o / 2)
    r = np.sqrt(1 - y ** 2)
    phi = ((i + rand) % n) * inc
    zero = np.zeros(n)
    return _f32(np.cos(phi) * r, y, np.sin(phi) * r, zero, zero, zero, np.full(n, 1.0 / n))


def lon_lat_spherical(n: int, seed: int = 0, R: float = 5.0) -> State:
    """LonLatSphericalUniverseGenerator.java:9-28."""
    g = rng(seed)
    lon = g.random(n) * 2 * math.pi
    lat = g.random(n) * 2 * math.pi
    zero = np.zeros(n)
    return _f32(R * np.cos(lat) * np.cos(lon), R * np.cos(lat) * np.sin(lon), R * np.sin(lat),
                zero, zero, zero, np.full(n, 1.0 / n))


def plummer_henon(n: int, seed: int = 0) -> State:
    """PlummerUniverseGenerator.java:6-43 (Plummer sphere in Henon units:
    radii scaled by 3 pi / 16, speeds by sqrt(16 / (3 pi)))."""
    g = rng(seed)
    rsc = (3 * math.pi) / 16
    vsc = math.sqrt(1.0 / rsc)
    X = np.maximum(g.random(n) * 0.999, 1e-300)
    r = 1.0 / np.sqrt(X ** (-2.0 / 3.0) - 1)
    p = _isotropic(g, n) * (rsc * r)[:, None]
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x = g.random(todo.size)
        y = g.random(todo.size) * 0.1
        ok = y <= x * x * (1 - x * x) ** 3.5
        q[todo[ok]] = x[ok]
        todo = todo[~ok]
    v = q * np.sqrt(2.0 / np.sqrt(1 + r * r))
    vel = _isotropic(g, n) * (vsc * v)[:, None]
    return _f32(p[:, 0], p[:, 1], p[:, 2], vel[:, 0], vel[:, 1], vel[:, 2], np.full(n, 1.0 / n))


def rotating_disk(n: int, radius: float = 3.5, velocity_multiplier: float = 1.0,
                  center_mass: float = 1.0, seed: int = 0) -> State:
    """RotatingDiskGalaxyGenerator.java:17-43: a central mass at the origin and
    a thin disk on circular orbits (FP32 arithmetic as in the reference)."""
    g = rng(seed)
    f = np.float32
    x = np.zeros(n, f); y = np.zeros(n, f); z = np.zeros(n, f)
    vx = np.zeros(n, f); vy = np.zeros(n, f); vz = np.zeros(n, f)
    m = np.zeros(n, f)
    m[0] = f(center_mass)
    k = n - 1
    r = (g.random(k) * radius).astype(f) + f(0.05)
    alpha = g.random(k) * 2 * math.pi
    xs = (np.cos(alpha) * r.astype(np.float64)).astype(f)
    ys = (np.sin(alpha) * r.astype(np.float64)).astype(f)
    x[1:], y[1:] = xs, ys
    z[1:] = ((g.random(k) - 0.5) / 8).astype(f)
    m[1:] = f(1.0) / f(n)
    v0 = np.sqrt(((m[0] + m[1:]) / (r * r * r)).astype(np.float64)).astype(f) * f(velocity_multiplier)
    vx[1:] = ys * v0
    vy[1:] = -xs * v0
    return x, y, z, vx, vy, vz, m


def two_body() -> State:
    """TwoBodyUniverse.java:8-18."""
    x = np.array([-1.1, 1], np.float32); y = np.array([-1, 1.1], np.float32)
    z = np.array([-1, 1], np.float32)
    zero = np.zeros(2, np.float32)
    return x, y, z, zero, zero.copy(), zero.copy(), np.full(2, 1 / 8, np.float32)


def eight_body() -> State:
    """EightBodyUniverse.java:8-48."""
    x = np.array([-1.1, 1, -1, -1, 1, -1, 1, 1], np.float32)
    y = np.array([-1, -1, 1, -1, 1, 1.1, -1, 1.1], np.float32)
    z = np.array([-1, -1, -1.1, 1.1, -1, 1, 1, 1], np.float32)
    zero = np.zeros(8, np.float32)
    return x, y, z, zero, zero.copy(), zero.copy(), np.full(8, 1 / 8, np.float32)


def big_tree() -> State:
    """BigTreeUniverse.java:8-24 (body 3's mass is never set there: 0)."""
    x = np.array([0, 1e-8, 10000, -10000], np.float32)
    y = np.array([0, 0, 10000, -10000], np.float32)
    z = np.array([0, 0, 10000, -10000], np.float32)
    zero = np.zeros(4, np.float32)
    return x, y, z, zero, zero.copy(), zero.copy(), np.array([0.1, 0.1, 0.1, 0.0], np.float32)


class ArrayUniverse:
    """UniverseGenerator over precomputed arrays (like SerializedUniverseGenerator,
    SerializedUniverseGenerator.java:39-53)."""

    def __init__(self, state: State):
        self.state = tuple(np.asarray(a, np.float32) for a in state)
        self.nbodies = len(self.state[0])

    def generate(self, offset, nbodies, bodiesX, bodiesY, bodiesZ, velX, velY, velZ, bodiesMass):
        if nbodies != self.nbodies:
            raise ValueError("invalid amount of bodies for this universe")
        for dst, src in zip((bodiesX, bodiesY, bodiesZ, velX, velY, velZ, bodiesMass), self.state):
            dst[offset:offset + nbodies] = src


class FunctionUniverse:
    """UniverseGenerator that calls a generator function of n."""

    def __init__(self, fn: Callable[..., State], **kw):
        self.fn, self.kw = fn, kw

    def generate(self, offset, nbodies, bodiesX, bodiesY, bodiesZ, velX, velY, velZ, bodiesMass):
        st = self.fn(nbodies, **self.kw)
        for dst, src in zip((bodiesX, bodiesY, bodiesZ, velX, velY, velZ, bodiesMass), st):
            dst[offset:offset + nbodies] = src


# ---- .universe files (UniverseSerializer.java:10-39, SerializedUniverseGenerator.java:21-53) ----
#
# The reference writes a universe with java.io.ObjectOutputStream: writeInt(N)
# (a 4-byte block-data record), then the seven float[] arrays x, y, z, vx, vy,
# vz, m as serialized objects (big-endian floats). An older layout found in the
# reference's repository root stores N as a boxed java.lang.Integer object and
# arrays longer than N (the first N entries are the bodies). Both are read here
# with a small reader of exactly these stream elements.

_STREAM_MAGIC = b"\xac\xed\x00\x05"
_TC_NULL, _TC_REFERENCE, _TC_CLASSDESC, _TC_OBJECT = 0x70, 0x71, 0x72, 0x73
_TC_ARRAY, _TC_BLOCKDATA, _TC_ENDBLOCKDATA = 0x75, 0x77, 0x78


class _JavaStream:
    def __init__(self, data: bytes):
        if data[:4] != _STREAM_MAGIC:
            raise ValueError("not a Java serialization stream")
        self.b, self.p = data, 4

    def u1(self) -> int:
        v = self.b[self.p]
        self.p += 1
        return v

    def take(self, n: int) -> bytes:
        if self.p + n > len(self.b):
            raise ValueError("truncated universe file")
        v = self.b[self.p:self.p + n]
        self.p += n
        return v

    def i4(self) -> int:
        return int.from_bytes(self.take(4), "big", signed=True)

    def classdesc(self) -> str:
        tc = self.u1()
        if tc == _TC_REFERENCE:
            self.take(4)
            return ""
        if tc == _TC_NULL:
            return ""
        if tc != _TC_CLASSDESC:
            raise ValueError(f"unexpected type code 0x{tc:02x} for a class descriptor")
        name = self.take(int.from_bytes(self.take(2), "big")).decode("utf-8")
        self.take(8)                      # serialVersionUID
        self.u1()                         # flags
        for _ in range(int.from_bytes(self.take(2), "big")):
            code = chr(self.u1())
            self.take(int.from_bytes(self.take(2), "big"))      # field name
            if code in "L[":              # object field: its type as a string
                if self.u1() == 0x74:     # TC_STRING
                    self.take(int.from_bytes(self.take(2), "big"))
                else:
                    self.take(4)          # TC_REFERENCE handle
        if self.u1() != _TC_ENDBLOCKDATA:
            raise ValueError("class annotations are not supported")
        self.classdesc()                  # superclass
        return name

    def count(self) -> int:
        tc = self.u1()
        if tc == _TC_BLOCKDATA:           # writeInt
            if self.u1() != 4:
                raise ValueError("unexpected block-data length")
            return self.i4()
        if tc == _TC_OBJECT:              # boxed Integer (older layout)
            self.classdesc()
            return self.i4()
        raise ValueError(f"unexpected type code 0x{tc:02x} for the body count")

    def float_array(self) -> np.ndarray:
        if self.u1() != _TC_ARRAY:
            raise ValueError("expected a float[] array")
        self.classdesc()
        n = self.i4()
        return np.frombuffer(self.take(4 * n), dtype=">f4").astype(np.float32)


def read_universe(path: str) -> State:
    """Bodies of a .universe file as float32 (x, y, z, vx, vy, vz, m)."""
    with open(path, "rb") as f:
        js = _JavaStream(f.read())
    n = js.count()
    arrays = [js.float_array() for _ in range(7)]
    if n < 0 or any(len(a) < n for a in arrays):
        raise ValueError("universe arrays are shorter than the body count")
    return tuple(np.ascontiguousarray(a[:n]) for a in arrays)


def write_universe(path: str, state: State) -> None:
    """Write bodies in the layout of UniverseSerializer.java:10-39 (writeInt(N),
    then the seven float[] arrays), readable by read_universe and by the
    reference's SerializedUniverseGenerator."""
    arrays = [np.asarray(a, np.float32) for a in state]
    n = len(arrays[0])
    out = bytearray(_STREAM_MAGIC)
    out += bytes([_TC_BLOCKDATA, 4]) + n.to_bytes(4, "big", signed=True)
    for k, a in enumerate(arrays):
        out.append(_TC_ARRAY)
        if k == 0:
            out += bytes([_TC_CLASSDESC]) + (2).to_bytes(2, "big") + b"[F"
            out += bytes.fromhex("0b9c818922e00c42")        # serialVersionUID of float[]
            out += bytes([0x02]) + (0).to_bytes(2, "big")   # SC_SERIALIZABLE, no fields
            out += bytes([_TC_ENDBLOCKDATA, _TC_NULL])
        else:
            out += bytes([_TC_REFERENCE]) + (0x7E0000).to_bytes(4, "big")   # handle of the float[] descriptor
        out += len(a).to_bytes(4, "big", signed=True) + a.astype(">f4").tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))


class FileUniverse(ArrayUniverse):
    """SerializedUniverseGenerator (SerializedUniverseGenerator.java:21-53): the
    bodies of a .universe file."""

    def __init__(self, path: str):
        super().__init__(read_universe(path))
