"""C ABI checks that need no GPU: libnbody.so loads, exports every function
declared in include/nbody.h, and rejects bad arguments before any HIP call."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nbody.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(nb_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    fns = declared_functions()
    for f in ("nb_create", "nb_destroy", "nb_set_state", "nb_step", "nb_copy_vertices",
              "nb_num_bodies", "nb_last_error"):
        assert f in fns


def test_library_exports_every_declared_symbol(nb):
    lib = nb.load()
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, f"declared in nbody.h but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", nb._lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (nb_\w+)", out))
    assert set(declared_functions()) <= exported


def test_library_is_gfx950(nb):
    """The device code object inside libnbody.so targets gfx950."""
    data = open(nb._lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_default_params_are_reference_constants(nb):
    p = nb.make_params(32768)
    assert p.n_bodies == 32768
    assert np.isclose(p.theta ** 2, 1.5)       # THETA (calculateforce.cl:16)
    assert np.isclose(p.softening ** 2, 0.0025)  # EPSILON (calculateforce.cl:14)
    assert np.isclose(p.dt, 0.025)             # TIMESTEP (calculateforce.cl:18)
    assert p.vote_width == 16                  # WARPSIZE (GPUBarnesHutNBodySimulation.java:47)
    assert p.pool_factor == 2.0                # NUMBER_OF_NODES = 2 N (java:220)


@pytest.mark.parametrize("field,value", [
    ("n_bodies", 0), ("vote_width", 17), ("vote_width", 8), ("softening", 0.0),
    ("softening", float("nan")), ("theta", -1.0), ("flags", 1 << 7), ("pool_factor", 0.5),
    ("mode", 5), ("n_bodies", 40_000_000),
])
def test_create_rejects_bad_params_without_gpu(nb, field, value):
    lib = nb.load()
    p = nb.make_params(1000)
    setattr(p, field, value)
    h = C.c_void_p()
    assert lib.nb_create(C.byref(p), C.byref(h)) == nb._lib.NB_E_ARG
    assert not h
    assert lib.nb_last_error(None) == b"invalid nb_params"


@pytest.mark.parametrize("kw", [
    dict(coarse_theta=0.3),                   # below theta
    dict(coarse_theta=1.0),                   # the extent test needs theta_c < 1
    dict(coarse_alpha=-0.01),                 # negative relative bound
    dict(coarse_alpha=float("inf")),
    dict(vote_width=16),                      # the coarse traversal is one wave64 per group
    dict(check_adapt=2),                      # 0 or 1
])
def test_learned_mode_parameter_checks(nb, kw):
    lib = nb.load()
    args = dict(theta=0.4, coarse_theta=0.95, coarse_alpha=0.02, vote_width=64, method=nb._lib.NB_MODE_LEARNED)
    args.update(kw)
    p = nb.make_params(1000, **args)
    h = C.c_void_p()
    assert lib.nb_create(C.byref(p), C.byref(h)) == nb._lib.NB_E_ARG


def test_default_learned_parameters(nb):
    p = nb.make_params(1000)
    assert np.isclose(p.coarse_theta, 0.95) and np.isclose(p.coarse_alpha, 0.02)
    assert p.check_adapt == 0  # a failed check keeps plain theta, as before the option existed


def test_null_handles(nb):
    lib = nb.load()
    E = nb._lib.NB_E_ARG
    assert lib.nb_step(None, 1) == E
    assert lib.nb_run_stage(None, 1) == E
    assert lib.nb_copy_vertices(None, None, None) == E
    assert lib.nb_num_bodies(None) == 0
    assert lib.nb_get_counters(None, None) == E
    assert lib.nb_get_control(None, None) == E
    assert lib.nb_set_control(None, None) == E
    assert lib.nb_create(None, None) == E
    lib.nb_destroy(None)


def test_control_record_layout(nb):
    """nb_control as the header lays it out: 8-byte fields first, 88 bytes."""
    import ctypes as C

    c = nb._lib.NbControl
    assert C.sizeof(c) == 88
    assert c.coarse_alpha.offset == 16 and c.fallbacks.offset == 32 and c.last_check_median.offset == 48


def test_status_strings(nb):
    lib = nb.load()
    for code in range(-8, 1):
        s = lib.nb_status_string(code)
        assert s and s != b"unknown status"
    assert lib.nb_status_string(-99) == b"unknown status"


def test_python_mirror_api(nb):
    """The Python mirror keeps the reference's plugin API names
    (AbstractNBodySimulation.java:16-43)."""
    from nbody_amd import AbstractNBodySimulation, HipBarnesHutNBodySimulation, Mode

    for name in ("step", "initGLBuffers", "getNumberOfBodies", "init"):
        assert name in AbstractNBodySimulation.__abstractmethods__
    assert {m.name for m in Mode} == {"GL_INTEROP", "DEFAULT"}
    assert issubclass(HipBarnesHutNBodySimulation, AbstractNBodySimulation)
