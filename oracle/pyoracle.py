"""ctypes binding of the CPU oracle (oracle/bh_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

FP64_POS = 1
COM_FP32 = 2

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.orc_nodes_for.argtypes = [C.c_int]
        L.orc_nodes_for.restype = C.c_int
        L.orc_create.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_set_state.argtypes = [vp] + [_f32p] * 7
        L.orc_set_accel.argtypes = [vp, _f32p, _f32p, _f32p, C.c_int]
        L.orc_set_positions64.argtypes = [vp, _f64p, _f64p, _f64p]
        L.orc_bbox_arrays.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p]
        L.orc_dq_table.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, _f32p]
        for name in ("orc_bbox", "orc_summarize", "orc_sort", "orc_integrate"):
            getattr(L, name).argtypes = [vp]
        for name in ("orc_build", "orc_force"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = C.c_int
        L.orc_force_groups.argtypes = [vp, _i32p, C.c_int]
        L.orc_force_groups.restype = C.c_int
        L.orc_force_fp32_groups.argtypes = [vp, _i32p, C.c_int, _f32p, _f32p, _f32p]
        L.orc_force_fp32_groups.restype = C.c_longlong
        L.orc_step.argtypes = [vp, C.c_int]
        L.orc_step.restype = C.c_int
        L.orc_run_stages.argtypes = [vp, C.c_int, C.c_int]
        L.orc_run_stages.restype = C.c_int
        L.orc_get_state.argtypes = [vp] + [_f32p] * 6
        L.orc_get_accel.argtypes = [vp] + [_f32p] * 3
        L.orc_get_positions64.argtypes = [vp] + [_f64p] * 3
        L.orc_get_sorted.argtypes = [vp, _i32p]
        L.orc_get_scalars.argtypes = [vp, _i64p]
        L.orc_get_root.argtypes = [vp, _f32p]
        L.orc_get_paths.argtypes = [vp, _u64p, _i32p]
        L.orc_get_paths.restype = C.c_int
        L.orc_get_cells.argtypes = [vp, _i32p, _f32p, _f32p, _f32p, _f32p, _i32p]
        L.orc_direct_fp64.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, C.c_double, C.c_int,
                                      _i32p, _f64p, _f64p, _f64p]
        L.orc_num_threads.restype = C.c_int
        L.orc_moments.argtypes = [vp]
        L.orc_get_moments.argtypes = [vp, C.c_int, _f64p]
        L.orc_get_moments.restype = C.c_int
        L.orc_coarse.argtypes = [vp, C.c_int, C.c_float, C.c_float, _i32p, C.c_int, vp, _f64p, vp]
        L.orc_coarse.restype = C.c_longlong
        L.orc_mlp_params.restype = C.c_int
        L.orc_mlp_forward.argtypes = [_f64p, _f32p, C.c_int, _f64p]
        L.orc_mlp_loss_grad.argtypes = [_f64p, _f32p, _f32p, C.c_int, _f64p]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def nodes_for(n: int) -> int:
    return lib().orc_nodes_for(n)


def bbox(x, y, z):
    """(root_x, root_y, root_z, radius, min_x, min_y, min_z) per boundingbox.cl."""
    out = np.zeros(7, np.float32)
    lib().orc_bbox_arrays(len(x), _f32(x), _f32(y), _f32(z), out)
    return out


def dq_table(radius, theta, softening, levels):
    out = np.zeros(max(levels, 1), np.float32)
    t2 = np.float32(theta) * np.float32(theta)
    e2 = np.float32(softening) * np.float32(softening)
    lib().orc_dq_table(float(radius), float(t2), float(e2), levels, out)
    return out


def direct_fp64(x, y, z, m, softening, idx):
    """FP64 direct-sum accelerations (G=1) on the bodies listed in idx. The
    squared softening is the FP32 value softening*softening, as the force
    kernels use it (calculateforce.cl:14: EPSILON is an FP32 constant)."""
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    ax = np.zeros(len(idx)); ay = np.zeros(len(idx)); az = np.zeros(len(idx))
    eps2 = float(np.float32(softening) * np.float32(softening))
    lib().orc_direct_fp64(len(x), _f32(x), _f32(y), _f32(z), _f32(m), eps2,
                          len(idx), idx, ax, ay, az)
    return np.stack([ax, ay, az], axis=1)


MLP_PARAMS = 5955


def mlp_forward(params, x):
    """FP64 MLP 24 -> 64 -> 64 -> 3 (tanh-GELU): y[n, 3] for x[n, 24]."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((len(x), 3), np.float64)
    lib().orc_mlp_forward(P, x.reshape(-1), len(x), y.reshape(-1))
    return y


def mlp_loss_grad(params, x, t):
    """(mean squared error over n*3 outputs, its gradient) in FP64."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(MLP_PARAMS, np.float64)
    loss = lib().orc_mlp_loss_grad(P, x.reshape(-1), t.reshape(-1), len(x), g)
    return loss, g


class Oracle:
    """One reference-semantics simulation on the CPU (mirrors the engine API)."""

    def __init__(self, n, theta=0.5, softening=0.05, dt=0.025, vote=16, flags=0):
        self.n = n
        self.vote = vote
        self.h = lib().orc_create(n, theta, softening, dt, vote, flags)
        if not self.h:
            raise ValueError("orc_create rejected the parameters")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, x, y, z, vx, vy, vz, m):
        lib().orc_set_state(self.h, _f32(x), _f32(y), _f32(z), _f32(vx), _f32(vy), _f32(vz), _f32(m))

    def set_accel(self, ax, ay, az, step):
        lib().orc_set_accel(self.h, _f32(ax), _f32(ay), _f32(az), int(step))

    def set_positions64(self, x, y, z):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        lib().orc_set_positions64(self.h, f(x), f(y), f(z))

    def step(self, k=1):
        return lib().orc_step(self.h, k)

    def run_stages(self, first, last):
        return lib().orc_run_stages(self.h, first, last)

    def force_groups(self, groups):
        g = np.ascontiguousarray(groups, dtype=np.int32)
        return lib().orc_force_groups(self.h, g, len(g))

    def force_fp32(self, groups=None):
        """The CPU baseline's FP32 force pass (cpu_fp32.c) over the listed vote
        groups (all when None), on this oracle's tree: (accelerations by sorted
        position [n, 3], rows of unlisted groups 0; interactions)."""
        if groups is None:
            groups = np.arange((self.n + self.vote - 1) // self.vote, dtype=np.int32)
        g = np.ascontiguousarray(groups, dtype=np.int32)
        a = [np.zeros(self.n, np.float32) for _ in range(3)]
        ni = lib().orc_force_fp32_groups(self.h, g, len(g), *a)
        if ni < 0:
            raise RuntimeError("orc_force_fp32_groups rejected the configuration")
        return np.stack(a, axis=1), int(ni)

    def state(self):
        arrs = [np.zeros(self.n, np.float32) for _ in range(6)]
        lib().orc_get_state(self.h, *arrs)
        return arrs

    def accel(self):
        arrs = [np.zeros(self.n, np.float32) for _ in range(3)]
        lib().orc_get_accel(self.h, *arrs)
        return np.stack(arrs, axis=1)

    def positions64(self):
        arrs = [np.zeros(self.n, np.float64) for _ in range(3)]
        lib().orc_get_positions64(self.h, *arrs)
        return np.stack(arrs, axis=1)

    def sorted(self):
        out = np.zeros(self.n, np.int32)
        lib().orc_get_sorted(self.h, out)
        return out

    def scalars(self):
        out = np.zeros(8, np.int64)
        lib().orc_get_scalars(self.h, out)
        keys = ("step", "bottom", "max_depth", "error", "cells", "nn", "interactions")
        return dict(zip(keys, (int(v) for v in out[:7])))

    def root(self):
        out = np.zeros(4, np.float32)
        lib().orc_get_root(self.h, out)
        return out

    def paths(self):
        """(paths uint64[n,3], depth int32[n]); valid after stage 2."""
        p = np.zeros((self.n, 3), np.uint64)
        d = np.zeros(self.n, np.int32)
        r = lib().orc_get_paths(self.h, p.reshape(-1), d)
        if r != self.n:
            raise RuntimeError(f"oracle tree reaches {r} of {self.n} bodies")
        return p, d

    def cells(self):
        sc = self.scalars()
        c = sc["cells"]
        cnt = np.zeros(c, np.int32)
        x = np.zeros(c, np.float32); y = np.zeros(c, np.float32); z = np.zeros(c, np.float32)
        m = np.zeros(c, np.float32)
        ch = np.zeros(8 * c, np.int32)
        lib().orc_get_cells(self.h, cnt, x, y, z, m, ch)
        return dict(count=cnt, x=x, y=y, z=z, m=m, child=ch.reshape(c, 8))

    def moments(self):
        """Compute every cell's second moments (after stage 3)."""
        lib().orc_moments(self.h)

    def cell_moments(self, cell):
        """(Qxx, Qyy, Qzz, Qxy, Qxz, Qyz, T, extent) of a cell (after stage 3)."""
        out = np.zeros(8, np.float64)
        if lib().orc_get_moments(self.h, int(cell), out) != 0:
            raise ValueError("no moments for that node")
        return out

    def coarse(self, theta, groups=None, features=True, alpha=0.0, geometric=False):
        """Coarse traversal (monopole + softened quadrupole) over the listed vote
        groups (all groups in order when None). geometric=False: the extent test
        at `theta` and the relative test at `alpha` (the learned mode's coarse
        traversal); geometric=True: the force stage's opening table of `theta`.
        Returns (acc [rows, 3], features [rows, 24] or None, e2 [rows, 3] or
        None, interactions); row i * vote + l is lane l of the i-th group (sorted
        position when all)."""
        if groups is None:
            groups = np.arange((self.n + self.vote - 1) // self.vote, dtype=np.int32)
        g = np.ascontiguousarray(groups, dtype=np.int32)
        rows = len(g) * self.vote
        acc = np.zeros((rows, 3), np.float64)
        feat = np.zeros((rows, 24), np.float32) if features else None
        e2 = np.zeros((rows, 3), np.float64) if features else None
        ni = lib().orc_coarse(self.h, 0 if geometric else 1, float(theta), float(alpha), g, len(g),
                              None if feat is None else feat.ctypes.data_as(C.c_void_p), acc.reshape(-1),
                              None if e2 is None else e2.ctypes.data_as(C.c_void_p))
        if ni < 0:
            raise RuntimeError("orc_coarse rejected the configuration")
        return acc, feat, e2, int(ni)
