"""The Java caller's call sequence (HipBarnesHutNBodySimulation.java through
nbody_jni.c), replayed from C through the glue's plain-C core
(gpu-nbody_amd/java/nbody_jni_core.c) by tests/jni_sequence.c against
libnbody.so -- this image has no JDK, so the JNI wrappers themselves are not
compiled here.

CPU: the glue's argument checks (array sizes, null handles, parameter packing
equal to nb_default_params for the Java defaults) run without any HIP call.
GPU: the whole sequence on the visualizer's default universe
(RotatingDiskGalaxyGenerator(3.5f, 1, 1) at 2048 * 16 bodies,
NBodyVisualizer.java:212): init, three GL_INTEROP frames whose vertices equal
the state bit for bit, the glue's rejections on a live handle, counters,
dispose; then the learned mode (train, model hand-over, four steps identical
on two handles, control state round trip, sampled force error).
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "bin", "jni_sequence")


def _bin():
    if not os.path.exists(BIN):
        raise FileNotFoundError(f"{BIN} is missing: run `python gpu-nbody_amd/build.py`")
    return BIN


def test_jni_glue_argument_checks(nb):
    r = subprocess.run([_bin(), "args"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "JNI_SEQUENCE_OK" in r.stdout


@pytest.mark.gpu
def test_jni_call_sequence(nb, tmp_path):
    from nbody_amd import universe as U

    n = 2048 * 16
    st = U.rotating_disk(n, 3.5, 1.0, 1.0, seed=5)
    path = tmp_path / "disk.bin"
    with open(path, "wb") as f:
        f.write(np.int32(n).tobytes())
        for a in st:
            f.write(np.ascontiguousarray(a, dtype="<f4").tobytes())
    r = subprocess.run([_bin(), "gpu", str(path)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "JNI_SEQUENCE_OK" in r.stdout
