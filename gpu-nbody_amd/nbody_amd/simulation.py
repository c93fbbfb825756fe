"""Host-side mirror of the reference's simulation plugin API.

AbstractNBodySimulation, its Mode enum and the UniverseGenerator interface keep
the reference's names and call order (AbstractNBodySimulation.java:5-44,
UniverseGenerator.java:7-23); HipBarnesHutNBodySimulation takes the place of
GPUBarnesHutNBodySimulation (GPUBarnesHutNBodySimulation.java:38-448) and
drives libnbody.so through its C ABI. As in the reference, where OpenCL API
errors throw (java:40-42), every failing call raises NbError.
"""
from __future__ import annotations

import abc
import ctypes as C
import enum
from typing import Optional, Protocol

import numpy as np

from . import _lib
from ._lib import NbCounters, NbError, NbParams, NbTrainParams, NbTrainReport, ptr


class Mode(enum.Enum):
    """AbstractNBodySimulation.Mode (AbstractNBodySimulation.java:41-43)."""

    GL_INTEROP = 0
    DEFAULT = 1


class UniverseGenerator(Protocol):
    """UniverseGenerator.generate (UniverseGenerator.java:22): fill bodies
    [offset, offset + nbodies) of the given arrays."""

    def generate(self, offset: int, nbodies: int, bodiesX: np.ndarray, bodiesY: np.ndarray,
                 bodiesZ: np.ndarray, velX: np.ndarray, velY: np.ndarray, velZ: np.ndarray,
                 bodiesMass: np.ndarray) -> None: ...


class AbstractNBodySimulation(abc.ABC):
    """AbstractNBodySimulation (AbstractNBodySimulation.java:5-44)."""

    def __init__(self, mode: Mode):
        self.mode = mode

    @abc.abstractmethod
    def step(self) -> None:
        """Executes the next step of the simulation."""

    @abc.abstractmethod
    def initGLBuffers(self, gl, positionVBO: int, velocityVBO: int) -> None:
        """Initializes the position and velocity buffers for the bodies."""

    @abc.abstractmethod
    def getNumberOfBodies(self) -> int:
        """Returns the number of bodies for this simulation."""

    @abc.abstractmethod
    def init(self, gl) -> None:
        """Initializes the simulation."""


def _check(lib, handle, code):
    if code != _lib.NB_OK:
        msg = lib.nb_last_error(handle)
        raise NbError(code, (msg.decode() if msg else "") or lib.nb_status_string(code).decode())
    return code


def make_params(nbodies: int, *, theta: Optional[float] = None, softening: Optional[float] = None,
                dt: Optional[float] = None, vote_width: Optional[int] = None, flags: int = 0,
                device: int = 0, pool_factor: Optional[float] = None, method: int = _lib.NB_MODE_PLAIN,
                coarse_theta: Optional[float] = None, check_tol: Optional[float] = None,
                check_every: Optional[int] = None, check_samples: Optional[int] = None,
                coarse_alpha: Optional[float] = None, check_adapt: Optional[int] = None) -> NbParams:
    """nb_params with the reference's defaults (nb_default_params) and overrides."""
    lib = _lib.load()
    p = NbParams()
    lib.nb_default_params(C.byref(p), int(nbodies))
    for name, val in (("theta", theta), ("softening", softening), ("dt", dt),
                      ("vote_width", vote_width), ("pool_factor", pool_factor),
                      ("coarse_theta", coarse_theta), ("check_tol", check_tol),
                      ("check_every", check_every), ("check_samples", check_samples),
                      ("coarse_alpha", coarse_alpha), ("check_adapt", check_adapt)):
        if val is not None:
            setattr(p, name, val)
    p.flags = int(flags)
    p.device = int(device)
    p.mode = int(method)
    return p


class HipBarnesHutNBodySimulation(AbstractNBodySimulation):
    """Barnes-Hut simulation on one MI355X through libnbody.so.

    Constructor arguments follow GPUBarnesHutNBodySimulation(mode, nbodies,
    generator) (java:104-108); the physics constants that the reference
    compiles into its kernels (calculateforce.cl:11-18) are keyword arguments
    here, defaulting to the same values.
    """

    def __init__(self, mode: Mode, nbodies: int, generator: UniverseGenerator, **params):
        super().__init__(mode)
        self.nbodies = int(nbodies)
        self.universeGenerator = generator
        self._params = make_params(self.nbodies, **params)
        self._lib = _lib.load()
        self._h = C.c_void_p()
        self._gl = None
        self.positionVBO = -1
        self.velocityVBO = -1
        self.positions: Optional[np.ndarray] = None
        self.velocities: Optional[np.ndarray] = None

    # -- AbstractNBodySimulation ------------------------------------------------
    def init(self, gl=None) -> None:
        """GPUBarnesHutNBodySimulation.init (java:111-151): create the device
        state and upload the generated universe."""
        lib = self._lib
        code = lib.nb_create(C.byref(self._params), C.byref(self._h))
        if code != _lib.NB_OK:
            msg = lib.nb_last_error(None)
            raise NbError(code, (msg.decode() if msg else "") or lib.nb_status_string(code).decode())
        n = self.nbodies
        arrs = [np.zeros(n, np.float32) for _ in range(7)]
        self.universeGenerator.generate(0, n, *arrs)
        self.set_state(*arrs)

    def initGLBuffers(self, gl, positionVBO: int, velocityVBO: int) -> None:
        """GPUBarnesHutNBodySimulation.initGLBuffers (java:229-246). There is no
        CL-GL sharing on the GPU side: the simulation keeps host float4 copies of
        the vertices and, in GL_INTEROP mode, uploads them into the given VBOs
        after every step (gl.glBufferSubData)."""
        self._gl = gl
        self.positionVBO = positionVBO
        self.velocityVBO = velocityVBO
        n = self.nbodies
        self.positions = np.zeros((n, 4), np.float32)
        self.velocities = np.zeros((n, 4), np.float32)

    def step(self) -> None:
        """GPUBarnesHutNBodySimulation.step (java:249-271)."""
        _check(self._lib, self._h, self._lib.nb_step(self._h, 1))
        if self.mode == Mode.GL_INTEROP and self.positions is not None:
            self.copy_vertices(self.positions, self.velocities)
            if self._gl is not None:
                self._upload(self._gl)

    def getNumberOfBodies(self) -> int:
        return self.nbodies

    # -- engine extras ----------------------------------------------------------
    def _upload(self, gl) -> None:
        GL_ARRAY_BUFFER = 0x8892
        gl.glBindBuffer(GL_ARRAY_BUFFER, self.positionVBO)
        gl.glBufferSubData(GL_ARRAY_BUFFER, 0, self.positions.nbytes, self.positions)
        gl.glBindBuffer(GL_ARRAY_BUFFER, self.velocityVBO)
        gl.glBufferSubData(GL_ARRAY_BUFFER, 0, self.velocities.nbytes, self.velocities)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            self._lib.nb_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, x, y, z, vx, vy, vz, m) -> None:
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        _check(self._lib, self._h, self._lib.nb_set_state(self._h, f(x), f(y), f(z), f(vx), f(vy), f(vz), f(m)))

    def steps(self, k: int) -> None:
        _check(self._lib, self._h, self._lib.nb_step(self._h, int(k)))

    def run_stage(self, stage: int) -> None:
        _check(self._lib, self._h, self._lib.nb_run_stage(self._h, int(stage)))

    def copy_vertices(self, pos4: Optional[np.ndarray], vel4: Optional[np.ndarray]) -> None:
        _check(self._lib, self._h, self._lib.nb_copy_vertices(self._h, ptr(pos4), ptr(vel4)))

    def copy_vertices_noalloc(self, pos4: Optional[np.ndarray], vel4: Optional[np.ndarray]) -> None:
        """copy_vertices without any device allocation (a staging array the
        library holds per GPU)."""
        _check(self._lib, self._h, self._lib.nb_copy_vertices_noalloc(self._h, ptr(pos4), ptr(vel4)))

    def state(self):
        n = self.nbodies
        a = [np.zeros(n, np.float32) for _ in range(6)]
        _check(self._lib, self._h, self._lib.nb_get_state(self._h, *[ptr(x) for x in a]))
        return a

    def accel(self) -> np.ndarray:
        n = self.nbodies
        a = [np.zeros(n, np.float32) for _ in range(3)]
        _check(self._lib, self._h, self._lib.nb_get_accel(self._h, *[ptr(x) for x in a]))
        return np.stack(a, axis=1)

    def positions64(self) -> np.ndarray:
        n = self.nbodies
        a = [np.zeros(n, np.float64) for _ in range(3)]
        _check(self._lib, self._h, self._lib.nb_get_positions64(self._h, *[ptr(x) for x in a]))
        return np.stack(a, axis=1)

    def root(self) -> np.ndarray:
        out = np.zeros(4, np.float32)
        _check(self._lib, self._h, self._lib.nb_get_root(self._h, out))
        return out

    def sorted(self) -> np.ndarray:
        out = np.zeros(self.nbodies, np.int32)
        _check(self._lib, self._h, self._lib.nb_get_sorted(self._h, out))
        return out

    def paths(self):
        p = np.zeros((self.nbodies, 3), np.uint64)
        d = np.zeros(self.nbodies, np.int32)
        _check(self._lib, self._h, self._lib.nb_get_paths(self._h, p.reshape(-1), d))
        return p, d

    def cells(self):
        nc = C.c_int32(0)
        cap = int(max(self.counters().live_cells, 1))
        cnt = np.zeros(cap, np.int32)
        xs = [np.zeros(cap, np.float32) for _ in range(4)]
        _check(self._lib, self._h, self._lib.nb_get_cells(self._h, cap, C.byref(nc), ptr(cnt),
                                                          *[ptr(a) for a in xs]))
        k = nc.value
        return dict(count=cnt[:k], x=xs[0][:k], y=xs[1][:k], z=xs[2][:k], m=xs[3][:k])

    def counters(self) -> NbCounters:
        c = NbCounters()
        _check(self._lib, self._h, self._lib.nb_get_counters(self._h, C.byref(c)))
        return c

    def probe_accel(self) -> np.ndarray:
        """Accelerations of the configured force method at the current positions."""
        n = self.nbodies
        a = [np.zeros(n, np.float32) for _ in range(3)]
        _check(self._lib, self._h, self._lib.nb_probe_accel(self._h, *[ptr(x) for x in a]))
        return np.stack(a, axis=1)

    def direct_accel(self, idx) -> np.ndarray:
        """FP64 direct-sum accelerations of the bodies in idx (computed on the GPU)."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros((len(idx), 3), np.float64)
        _check(self._lib, self._h, self._lib.nb_direct_accel(self._h, idx, len(idx), ptr(out)))
        return out

    def force_error(self, nsamples: int = 4096, seed: int = 1000) -> dict:
        """Relative force error against the FP64 direct sum on sampled bodies."""
        st = np.zeros(4, np.float64)
        _check(self._lib, self._h, self._lib.nb_force_error(self._h, int(nsamples), int(seed), ptr(st)))
        return dict(median=st[0], p90=st[1], max=st[2], mean=st[3])

    def set_timing(self, on: bool) -> None:
        _check(self._lib, self._h, self._lib.nb_set_timing(self._h, 1 if on else 0))

    # -- learned coarse-theta correction -------------------------------------------
    def train(self, **overrides) -> dict:
        """nb_train: generate pairs over `steps` plain steps, train, install.
        Keyword overrides of nb_train_params (steps, groups_per_step, iters,
        batch, lr, lr_final, beta1, beta2, adam_eps, seed)."""
        tp = NbTrainParams()
        self._lib.nb_default_train_params(C.byref(tp))
        for k, v in overrides.items():
            if not hasattr(tp, k):
                raise TypeError(f"unknown training parameter {k}")
            setattr(tp, k, v)
        rep = NbTrainReport()
        _check(self._lib, self._h, self._lib.nb_train(self._h, C.byref(tp), C.byref(rep)))
        return dict(samples=rep.samples, loss_first=rep.loss_first, loss_last=rep.loss_last,
                    gen_ms=rep.gen_ms, train_ms=rep.train_ms)

    def set_model(self, blob) -> None:
        blob = np.ascontiguousarray(blob, dtype=np.float32)
        _check(self._lib, self._h, self._lib.nb_set_model(self._h, blob, len(blob)))

    def get_model(self) -> np.ndarray:
        out = np.zeros(_lib.NB_MODEL_FLOATS, np.float32)
        _check(self._lib, self._h, self._lib.nb_get_model(self._h, out, len(out)))
        return out

    def coarse_features(self):
        """(features [n, 24], coarse acceleration [n, 3], e2 [n, 3]) by body index."""
        n = self.nbodies
        f = np.zeros((n, _lib.NB_MLP_IN), np.float32)
        a = np.zeros((n, 3), np.float32)
        e2 = np.zeros((n, 3), np.float32)
        _check(self._lib, self._h, self._lib.nb_coarse_features(self._h, ptr(f), ptr(a), ptr(e2)))
        return f, a, e2

    def mlp_forward(self, x) -> np.ndarray:
        """Raw features [n, 24] -> residual [n, 3] through the installed network (BF16 matrix cores)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), 3), np.float32)
        _check(self._lib, self._h, self._lib.nb_mlp_forward(self._h, x.reshape(-1), len(x), y.reshape(-1)))
        return y

    def mlp_loss_grad(self, x, t):
        """FP32 training kernels: (loss, gradient) for normalized x [n, 24], t [n, 3]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(_lib.NB_MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(self._lib, self._h, self._lib.nb_mlp_loss_grad(self._h, x.reshape(-1), t.reshape(-1), len(x),
                                                              C.byref(loss), g))
        return float(loss.value), g

    def mlp_adam(self, x, t, iters: int, lr: float) -> np.ndarray:
        """`iters` full-batch Adam steps from the installed parameters; returns the parameters."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        p = np.zeros(_lib.NB_MLP_PARAMS, np.float32)
        _check(self._lib, self._h, self._lib.nb_mlp_adam(self._h, x.reshape(-1), t.reshape(-1), len(x),
                                                         int(iters), float(lr), p))
        return p

    # -- diagnostics and checkpoints ---------------------------------------------
    def diagnostics(self, pe_samples: int = 0, seed: int = 1) -> dict:
        """FP64 kinetic and softened potential energy (exact when pe_samples <= 0),
        momentum, angular momentum, centre of mass and total mass."""
        out = np.zeros(12, np.float64)
        _check(self._lib, self._h, self._lib.nb_diagnostics(self._h, int(pe_samples), int(seed), ptr(out)))
        return dict(kinetic=out[0], potential=out[1], energy=out[0] + out[1], momentum=out[2:5].copy(),
                    angular_momentum=out[5:8].copy(), centre_of_mass=out[8:11].copy(), mass=out[11])

    def masses(self) -> np.ndarray:
        m = np.zeros(self.nbodies, np.float32)
        _check(self._lib, self._h, self._lib.nb_get_masses(self._h, ptr(m)))
        return m

    def control(self) -> "_lib.NbControl":
        """Learned-mode control state (check schedule, sample seed, alpha in
        effect, fallback)."""
        c = _lib.NbControl()
        _check(self._lib, self._h, self._lib.nb_get_control(self._h, C.byref(c)))
        return c

    def set_control(self, c: "_lib.NbControl") -> None:
        _check(self._lib, self._h, self._lib.nb_set_control(self._h, C.byref(c)))

    def checkpoint(self, path: str) -> None:
        """Write the full dynamic state (npz): positions (FP64 when kept), velocities,
        last accelerations, masses, step counter and running maximum depth; in the
        learned mode also the model blob and the control state (nb_control)."""
        c = self.counters()
        x, y, z, vx, vy, vz = self.state()
        pos = (self.positions64() if self._params.flags & _lib.NB_FP64_POSITIONS
               else np.stack([x, y, z], 1).astype(np.float64))
        extra = {}
        if self._params.mode == _lib.NB_MODE_LEARNED:
            try:
                extra["model"] = self.get_model()
            except NbError as e:
                if e.code != _lib.NB_E_MODEL:
                    raise
            ctl = self.control()
            extra["control"] = np.frombuffer(bytes(ctl), np.uint8).copy()
        np.savez(path, pos=pos, vel=np.stack([vx, vy, vz], 1), acc=self.accel(), m=self.masses(),
                 step=np.int32(c.step), max_depth=np.int32(c.max_depth), **extra)

    def restore(self, path: str) -> None:
        """Resume from a checkpoint written by checkpoint()."""
        f = np.load(path)
        pos, vel, acc = (np.ascontiguousarray(f[k]) for k in ("pos", "vel", "acc"))
        cols = [np.ascontiguousarray(pos[:, k], dtype=np.float64) for k in range(3)]
        cols += [np.ascontiguousarray(vel[:, k], dtype=np.float32) for k in range(3)]
        cols += [np.ascontiguousarray(acc[:, k], dtype=np.float32) for k in range(3)]
        cols.append(np.ascontiguousarray(f["m"], dtype=np.float32))
        _check(self._lib, self._h, self._lib.nb_restore_state(self._h, *[ptr(a) for a in cols],
                                                              int(f["step"]), int(f["max_depth"])))
        if "model" in f.files:
            self.set_model(f["model"])
        if "control" in f.files:
            raw = f["control"].tobytes()
            if len(raw) != C.sizeof(_lib.NbControl):
                raise ValueError("checkpoint control record has the wrong size")
            self.set_control(_lib.NbControl.from_buffer_copy(raw))
