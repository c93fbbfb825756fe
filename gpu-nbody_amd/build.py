"""Build libnbody.so (HIP, gfx950) in-tree.

    python gpu-nbody_amd/build.py            # compile every source, link the library

Each .hip/.cpp file in csrc/ is compiled by hipcc for gfx950 and linked into
gpu-nbody_amd/libnbody.so. Packed FP32 instructions are disabled for all
device code (-packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnbody.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ARCH = os.environ.get("NB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *COMMON, "-x", "hip", *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


JAVA = os.path.join(HERE, "java")
ROOT = os.path.dirname(HERE)
SEQ_SRC = os.path.join(ROOT, "tests", "jni_sequence.c")
SEQ_BIN = os.path.join(ROOT, "tests", "bin", "jni_sequence")


def build_jni_sequence(verbose: bool = False) -> str:
    """Compile the JNI glue's plain-C core (java/nbody_jni_core.c) with the
    program that replays the Java class's call sequence through it
    (tests/jni_sequence.c), linked against libnbody.so. The JNI wrappers
    themselves (java/nbody_jni.c) need a JDK's jni.h, which this image lacks."""
    os.makedirs(os.path.dirname(SEQ_BIN), exist_ok=True)
    cc = shutil.which("gcc") or "cc"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    rpath = "-Wl,-rpath,$ORIGIN/../../gpu-nbody_amd"
    cmd = [cc, "-O2", "-std=c11", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{JAVA}", SEQ_SRC,
           os.path.join(JAVA, "nbody_jni_core.c"), f"-L{HERE}", "-lnbody", f"-L{rocm_lib}", rpath,
           f"-Wl,-rpath,{rocm_lib}", "-lm", "-o", SEQ_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"jni sequence build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {SEQ_BIN}")
    return SEQ_BIN


if __name__ == "__main__":
    try:
        build(verbose=True)
        build_jni_sequence(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
